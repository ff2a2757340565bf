// KPP-Standalone chemical-state files (host code): the file that
// GeosCore/kppsa_interface_mod.F90:569-843 (KppSa_Write_Samples) writes for
// one grid cell, and a reader for it.  The numbers use the reference's
// Fortran edit descriptors, rendered here by the Fortran rules:
//   ESw.dEe  one nonzero digit before the point, d after, exponent of e
//            digits (es25.16e3 for values, es10.2e2 for tolerances);
//   Ew.d     0.ddddE+xx, or 0.dddd+xxx for a three-digit exponent;
//   Fw.d, Iw plain fields;
// right-justified in w columns, w asterisks when the value does not fit,
// "Infinity"/"NaN" for non-finite values.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gckpp_hip.h"
#include "gckpp_internal.h"

namespace {

int kfail(const char *fmt, ...) {
  char msg[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  return gckpp::set_error(msg);
}

std::string stars(int w) { return std::string((size_t)w, '*'); }

std::string rjust(const std::string &s, int w) {
  if ((int)s.size() > w) return stars(w);
  return std::string((size_t)(w - (int)s.size()), ' ') + s;
}

bool nonfinite(double x, int w, std::string &out) {
  if (std::isnan(x)) { out = rjust("NaN", w); return true; }
  if (std::isinf(x)) { out = rjust(x < 0 ? "-Infinity" : "Infinity", w); return true; }
  return false;
}

// ESw.dEe
std::string fmt_es(double x, int w, int d, int e) {
  std::string out;
  if (nonfinite(x, w, out)) return out;
  char buf[128];
  snprintf(buf, sizeof buf, "%.*E", d, x);  // [-]d.dddE+x..
  char *ep = strchr(buf, 'E');
  const int ex = atoi(ep + 1);
  *ep = 0;
  const int ax = ex < 0 ? -ex : ex;
  char xb[16];
  snprintf(xb, sizeof xb, "%0*d", e, ax);
  if ((int)strlen(xb) > e) return stars(w);
  return rjust(std::string(buf) + "E" + (ex < 0 ? "-" : "+") + xb, w);
}

// Ew.d: 0.ddd...E+xx (|exponent| <= 99) or 0.ddd...+xxx (<= 999); the zero
// before the point is dropped when the field has no room for it.
std::string fmt_e(double x, int w, int d) {
  std::string out;
  if (nonfinite(x, w, out)) return out;
  std::string mant;
  int ex = 0;
  if (x == 0.0) {
    mant = "." + std::string((size_t)d, '0');
  } else {
    char buf[128];
    snprintf(buf, sizeof buf, "%.*E", d - 1, fabs(x));  // d.ddd E+x: value = 0.dddd * 10^(x+1)
    char *ep = strchr(buf, 'E');
    ex = atoi(ep + 1) + 1;
    *ep = 0;
    std::string digs;
    for (char *p = buf; *p; ++p)
      if (isdigit((unsigned char)*p)) digs += *p;
    mant = "." + digs;
  }
  const int ax = ex < 0 ? -ex : ex;
  char xb[16];
  if (ax <= 99) snprintf(xb, sizeof xb, "E%c%02d", ex < 0 ? '-' : '+', ax);
  else if (ax <= 999) snprintf(xb, sizeof xb, "%c%03d", ex < 0 ? '-' : '+', ax);
  else return stars(w);
  std::string body = mant + xb;
  const std::string sign = std::signbit(x) ? "-" : "";
  if ((int)(sign.size() + 1 + body.size()) <= w) return rjust(sign + "0" + body, w);
  return rjust(sign + body, w);
}

// Fw.d
std::string fmt_f(double x, int w, int d) {
  std::string out;
  if (nonfinite(x, w, out)) return out;
  char buf[512];
  snprintf(buf, sizeof buf, "%.*f", d, x);
  std::string s(buf);
  if (s.size() > (size_t)w && s[0] == '0') s = s.substr(1);       // optional leading zero
  if (s.size() > (size_t)w && s.rfind("-0.", 0) == 0) s = "-" + s.substr(2);
  return rjust(s, w);
}

std::string fmt_i(long long v, int w) { return rjust(std::to_string(v), w); }

// A name record of GCKPP_NAME_LEN bytes: up to the NUL, trailing blanks
// dropped.
std::string name_of(const char *rec) {
  size_t n = 0;
  while (n < GCKPP_NAME_LEN && rec[n]) ++n;
  while (n > 0 && rec[n - 1] == ' ') --n;
  size_t b = 0;
  while (b < n && rec[b] == ' ') ++b;
  return std::string(rec + b, n - b);
}

// The header: 60 lines (the first gives the count), the column line after.
// Descriptive lines are this project's; the metadata keys and their
// formats, the line positions and the count are the file format's.
const char *const kIntro[] = {
    "",
    "                  KPP Standalone Atmospheric Chemical State",
    "File Description:",
    "One grid cell's full chemical state before the KPP Integrate call of a",
    "GEOS-Chem chemistry step, in the form a KPP Standalone program reads to",
    "repeat that cell's chemistry on its own.",
    "The KPP Standalone uses the concentrations, the rate constants and the",
    "integrator fields below; the other fields describe the cell.",
    "The first line of this file is the number of header lines; the CSV",
    "section follows the header.",
    "Units: molec/cm3 for concentrations, molec/cm3/s for the reaction",
    "rates (A), the KPP units of each rate constant (R).",
    "Fixed (non-integrated) species follow the variable ones.",
    "",
    "Written by the gckpp_hip library",
    "   (the MI355X chemistry solve of GEOS-Chem's KPP mechanisms)",
    "in the format of the GEOS-Chem KPP Standalone interface",
    "GeosCore/kppsa_interface_mod.F90",
    "     Mechanisms:",
    "        fullchem",
    "        Hg",
    "        carbon",
    "",
    "",
};
}  // namespace

extern "C" int gckpp_kppsa_write(int mech_id, const char *path, const gckpp_kppsa_meta *m,
                                 const char *names, const double *conc, const double *atol,
                                 const double *rconst, const double *aout) {
  if (mech_id < 0 || mech_id > 2) return kfail("gckpp_kppsa_write: unknown mechanism id %d", mech_id);
  if (!path || !m || !conc || !atol || !rconst || !aout)
    return kfail("gckpp_kppsa_write: null argument");
  const gckpp::MechInfo mi = gckpp::mech_info(mech_id);
  std::vector<std::string> L;
  L.push_back("60");
  L.push_back(std::string(76, '='));
  for (const char *t : kIntro) L.push_back(t);
  // 26 lines so far
  L.push_back("Meteorological and general grid cell metadata    ");
  L.push_back("");
  const std::string loc(m->location, strnlen(m->location, sizeof m->location));
  const std::string ts(m->timestamp, strnlen(m->timestamp, sizeof m->timestamp));
  L.push_back("Location:                                        " + loc);
  L.push_back("Timestamp:                                       " + ts);
  L.push_back("Longitude (degrees):                             " + fmt_f(m->lon, 11, 4));
  L.push_back("Latitude (degrees):                              " + fmt_f(m->lat, 11, 4));
  L.push_back("GEOS-Chem Vertical Level:                        " + fmt_i(m->level, 6));
  L.push_back("Pressure (hPa):                                  " + fmt_f(m->pressure_hpa, 11, 4));
  L.push_back("Temperature (K):                                 " + fmt_f(m->temperature, 11, 2));
  L.push_back("Dry air density (molec/cm3):                     " + fmt_e(m->airnumden, 11, 4));
  L.push_back("Water vapor mixing ratio (vol H2O/vol dry air):  " + fmt_e(m->h2o_vmr, 11, 4));
  L.push_back("Cloud fraction:                                  " + fmt_e(m->cloud_frac, 11, 4));
  L.push_back("Cosine of solar zenith angle:                    " + fmt_e(m->suncos, 11, 4));
  L.push_back("");
  L.push_back("KPP Integrator-specific parameters               ");
  L.push_back("");
  L.push_back("Init KPP Timestep (seconds):                     " + fmt_f(m->h_init, 11, 4));
  L.push_back("Exit KPP Timestep (seconds):                     " + fmt_f(m->h_exit, 11, 4));
  L.push_back("Chemistry operator timestep (seconds):           " + fmt_f(m->dt, 11, 4));
  L.push_back("Number of internal timesteps:                    " + fmt_i(m->nsteps, 6));
  L.push_back("ICNTRL integrator options used:");
  for (int r = 0; r < 2; ++r) {
    std::string s;
    for (int k = 0; k < 10; ++k) s += fmt_i(m->icntrl[10 * r + k], 6);
    L.push_back(s);
  }
  L.push_back("RCNTRL integrator options used:");
  for (int r = 0; r < 4; ++r) {
    std::string s;
    for (int k = 0; k < 5; ++k) s += fmt_f(m->rcntrl[5 * r + k], 13, 6);
    L.push_back(s);
  }
  L.push_back("");
  L.push_back("CSV data of full chemical state, including species concentrations,");
  L.push_back("rate constants (R) and instantaneous reaction rates (A).");
  L.push_back("All concentration units are in molec/cm3 and rates in molec/cm3/s.");
  L.push_back("");
  L.push_back(std::string(76, '='));
  if (L.size() != 60) return kfail("gckpp_kppsa_write: header has %zu lines", L.size());
  L.push_back("Name,   Value,   Absolute Tolerance");
  FILE *f = fopen(path, "w");
  if (!f) return kfail("gckpp_kppsa_write: cannot open %s: %s", path, strerror(errno));
  for (const std::string &s : L) fprintf(f, "%s\n", s.c_str());
  for (int n = 0; n < mi.nspec; ++n) {
    std::string nm = names ? name_of(names + (size_t)n * GCKPP_NAME_LEN) : std::string();
    if (nm.empty()) nm = "C" + std::to_string(n + 1);
    fprintf(f, "%s,%s,%s\n", nm.c_str(), fmt_es(conc[n], 25, 16, 3).c_str(),
            fmt_es(atol[n], 10, 2, 2).c_str());
  }
  for (int n = 0; n < mi.nreact; ++n) fprintf(f, "R%d,%s\n", n + 1, fmt_es(rconst[n], 25, 16, 3).c_str());
  for (int n = 0; n < mi.nreact; ++n) fprintf(f, "A%d,%s\n", n + 1, fmt_es(aout[n], 25, 16, 3).c_str());
  const bool bad = ferror(f) != 0;
  if (fclose(f) != 0 || bad) return kfail("gckpp_kppsa_write: writing %s failed", path);
  gckpp::set_error("");
  return 0;
}

namespace {

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && isspace((unsigned char)s[a])) ++a;
  while (b > a && isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

// A number in any of the formats above ("0.1234+100": an Ew.d three-digit
// exponent without its E).
bool parse_num(const std::string &t0, double &v) {
  std::string t = trim(t0);
  if (t.empty() || t[0] == '*') return false;
  for (size_t i = 1; i < t.size(); ++i)
    if ((t[i] == '+' || t[i] == '-') && isdigit((unsigned char)t[i - 1])) {
      t.insert(i, "E");
      break;
    }
  char *end = nullptr;
  v = strtod(t.c_str(), &end);
  return end && *end == 0;
}

bool after_key(const std::string &line, const char *key, std::string &rest) {
  const size_t n = strlen(key);
  if (line.compare(0, n, key) != 0) return false;
  rest = line.substr(n);
  return true;
}

}  // namespace

extern "C" int gckpp_kppsa_read(int mech_id, const char *path, gckpp_kppsa_meta *m, char *names,
                                double *conc, double *atol, double *rconst, double *aout) {
  if (mech_id < 0 || mech_id > 2) return kfail("gckpp_kppsa_read: unknown mechanism id %d", mech_id);
  if (!path) return kfail("gckpp_kppsa_read: null path");
  const gckpp::MechInfo mi = gckpp::mech_info(mech_id);
  FILE *f = fopen(path, "r");
  if (!f) return kfail("gckpp_kppsa_read: cannot open %s: %s", path, strerror(errno));
  std::vector<std::string> lines;
  {
    char buf[4096];
    std::string cur;
    while (fgets(buf, sizeof buf, f)) {
      cur += buf;
      if (!cur.empty() && cur.back() == '\n') {
        cur.pop_back();
        if (!cur.empty() && cur.back() == '\r') cur.pop_back();
        lines.push_back(cur);
        cur.clear();
      }
    }
    if (!cur.empty()) lines.push_back(cur);
    fclose(f);
  }
  if (lines.empty()) return kfail("gckpp_kppsa_read: %s is empty", path);
  char *endp = nullptr;
  const long nhead = strtol(lines[0].c_str(), &endp, 10);
  if (endp == lines[0].c_str() || nhead < 1 || (size_t)nhead >= lines.size())
    return kfail("gckpp_kppsa_read: %s: first line is not a header line count", path);
  gckpp_kppsa_meta mm;
  memset(&mm, 0, sizeof mm);
  int icrow = -1, rcrow = -1;
  for (long i = 1; i < nhead; ++i) {
    const std::string &s = lines[(size_t)i];
    std::string r;
    double v;
    if (icrow >= 0 && icrow < 2) {
      for (int k = 0; k < 10; ++k)
        mm.icntrl[10 * icrow + k] = (int32_t)strtol(s.substr((size_t)6 * k, 6).c_str(), nullptr, 10);
      ++icrow;
      continue;
    }
    if (rcrow >= 0 && rcrow < 4) {
      for (int k = 0; k < 5; ++k)
        if (parse_num(s.substr((size_t)13 * k, 13), v)) mm.rcntrl[5 * rcrow + k] = v;
      ++rcrow;
      continue;
    }
    if (after_key(s, "Location:", r)) {
      snprintf(mm.location, sizeof mm.location, "%s", trim(r).c_str());
    } else if (after_key(s, "Timestamp:", r)) {
      snprintf(mm.timestamp, sizeof mm.timestamp, "%s", trim(r).c_str());
    } else if (after_key(s, "Longitude (degrees):", r)) {
      if (parse_num(r, v)) mm.lon = v;
    } else if (after_key(s, "Latitude (degrees):", r)) {
      if (parse_num(r, v)) mm.lat = v;
    } else if (after_key(s, "GEOS-Chem Vertical Level:", r)) {
      mm.level = (int32_t)strtol(r.c_str(), nullptr, 10);
    } else if (after_key(s, "Pressure (hPa):", r)) {
      if (parse_num(r, v)) mm.pressure_hpa = v;
    } else if (after_key(s, "Temperature (K):", r)) {
      if (parse_num(r, v)) mm.temperature = v;
    } else if (after_key(s, "Dry air density (molec/cm3):", r)) {
      if (parse_num(r, v)) mm.airnumden = v;
    } else if (after_key(s, "Water vapor mixing ratio (vol H2O/vol dry air):", r)) {
      if (parse_num(r, v)) mm.h2o_vmr = v;
    } else if (after_key(s, "Cloud fraction:", r)) {
      if (parse_num(r, v)) mm.cloud_frac = v;
    } else if (after_key(s, "Cosine of solar zenith angle:", r)) {
      if (parse_num(r, v)) mm.suncos = v;
    } else if (after_key(s, "Init KPP Timestep (seconds):", r)) {
      if (parse_num(r, v)) mm.h_init = v;
    } else if (after_key(s, "Exit KPP Timestep (seconds):", r)) {
      if (parse_num(r, v)) mm.h_exit = v;
    } else if (after_key(s, "Chemistry operator timestep (seconds):", r)) {
      if (parse_num(r, v)) mm.dt = v;
    } else if (after_key(s, "Number of internal timesteps:", r)) {
      mm.nsteps = (int32_t)strtol(r.c_str(), nullptr, 10);
    } else if (after_key(s, "ICNTRL integrator options used:", r)) {
      icrow = 0;
    } else if (after_key(s, "RCNTRL integrator options used:", r)) {
      rcrow = 0;
    }
  }
  // the column line, then NSPEC species, NREACT R, NREACT A lines
  size_t i = (size_t)nhead;
  if (lines[i].compare(0, 5, "Name,") == 0) ++i;
  const size_t need = (size_t)mi.nspec + 2 * (size_t)mi.nreact;
  if (lines.size() - i < need)
    return kfail("gckpp_kppsa_read: %s has %zu data lines, the mechanism needs %zu", path,
                 lines.size() - i, need);
  for (int n = 0; n < mi.nspec; ++n, ++i) {
    const std::string &s = lines[i];
    const size_t c1 = s.find(','), c2 = c1 == std::string::npos ? c1 : s.find(',', c1 + 1);
    double v, a;
    if (c2 == std::string::npos || !parse_num(s.substr(c1 + 1, c2 - c1 - 1), v) ||
        !parse_num(s.substr(c2 + 1), a))
      return kfail("gckpp_kppsa_read: %s line %zu: not a species line: %s", path, i + 1, s.c_str());
    if (conc) conc[n] = v;
    if (atol) atol[n] = a;
    if (names) {
      char *rec = names + (size_t)n * GCKPP_NAME_LEN;
      memset(rec, 0, GCKPP_NAME_LEN);
      snprintf(rec, GCKPP_NAME_LEN, "%s", trim(s.substr(0, c1)).c_str());
    }
  }
  for (int part = 0; part < 2; ++part) {
    const char tag = part == 0 ? 'R' : 'A';
    double *out = part == 0 ? rconst : aout;
    for (int n = 0; n < mi.nreact; ++n, ++i) {
      const std::string &s = lines[i];
      const size_t c1 = s.find(',');
      double v;
      if (s.empty() || s[0] != tag || c1 == std::string::npos ||
          strtol(s.c_str() + 1, nullptr, 10) != n + 1 || !parse_num(s.substr(c1 + 1), v))
        return kfail("gckpp_kppsa_read: %s line %zu: expected %c%d: %s", path, i + 1, tag, n + 1,
                     s.c_str());
      if (out) out[n] = v;
    }
  }
  if (m) *m = mm;
  gckpp::set_error("");
  return 0;
}
