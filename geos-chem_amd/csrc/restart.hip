// The KPPHvalue restart variable (host code): GEOS-Chem keeps each cell's
// last predicted step size, RSTATUS(Nhnew), in State_Chm%KPPHvalue
// (Headers/state_chm_mod.F90:1742-1753) and reads it back at start-up from
// the restart file's Chem_KPPHvalue(lon,lat,lev,time), zero when the file
// has none (GeosCore/hco_utilities_gc_mod.F90:2144-2177).  This writes and
// reads that variable in a netCDF classic file (the big-endian format of
// the netCDF-3 specification, CDF-1 or CDF-2), with no netCDF library.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gckpp_hip.h"
#include "gckpp_internal.h"

namespace {

constexpr int NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6, NC_BYTE = 1;
constexpr uint32_t NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C;
const char *const kVar = "Chem_KPPHvalue";
const char *const kTagAtt = "gckpp_step_model";

int rfail(const char *fmt, const char *a = "", const char *b = "") {
  char msg[512];
  snprintf(msg, sizeof msg, fmt, a, b);
  return gckpp::set_error(msg);
}

struct Out {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int k = 3; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
  }
  void u64(uint64_t v) {
    for (int k = 7; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
  }
  void pad4() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad4();
  }
  void text_att(const std::string &n, const std::string &v) {
    name(n);
    u32(NC_CHAR);
    u32((uint32_t)v.size());
    b.insert(b.end(), v.begin(), v.end());
    pad4();
  }
  void f64(double x) {
    uint64_t v;
    memcpy(&v, &x, 8);
    u64(v);
  }
};

struct In {
  const std::vector<unsigned char> &b;
  size_t p = 0;
  bool ok = true;
  explicit In(const std::vector<unsigned char> &v) : b(v) {}
  uint32_t u32() {
    if (p + 4 > b.size()) { ok = false; return 0; }
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) v = (v << 8) | b[p++];
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return (hi << 32) | u32();
  }
  std::string name() {
    const uint32_t n = u32();
    if (!ok || p + n > b.size()) { ok = false; return std::string(); }
    std::string s(b.begin() + (long)p, b.begin() + (long)(p + n));
    p += (n + 3) / 4 * 4;
    return s;
  }
};

int type_size(int t) {
  switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
  }
  return 0;
}

// Skips (or, for char attributes named want, returns) an attribute list.
bool att_list(In &r, const char *want, std::string *val) {
  const uint32_t tag = r.u32(), n = r.u32();
  if (!r.ok || (tag != 0 && tag != NC_ATTRIBUTE)) return false;
  for (uint32_t i = 0; i < n && r.ok; ++i) {
    const std::string nm = r.name();
    const int t = (int)r.u32();
    const uint32_t ne = r.u32();
    const int ts = type_size(t);
    if (!r.ok || ts == 0) return false;
    const size_t bytes = (size_t)ne * ts;
    if (r.p + bytes > r.b.size()) return false;
    if (want && val && t == NC_CHAR && nm == want)
      val->assign(r.b.begin() + (long)r.p, r.b.begin() + (long)(r.p + bytes));
    r.p += (bytes + 3) / 4 * 4;
  }
  return r.ok;
}

}  // namespace

extern "C" int gckpp_kpph_restart_write(const char *path, int64_t nx, int64_t ny, int64_t nz,
                                        const double *kpph, const char *model_tag) {
  if (!path || !kpph) return rfail("gckpp_kpph_restart_write: null argument");
  if (nx <= 0 || ny <= 0 || nz <= 0 || nx * ny * nz * 8 >= (int64_t)0xFFFFFFFFLL)
    return rfail("gckpp_kpph_restart_write: grid sizes out of range");
  const uint64_t n = (uint64_t)(nx * ny * nz);
  Out o;
  o.b = {'C', 'D', 'F', 2};
  o.u32(0);                                   // numrecs (no record variable)
  o.u32(NC_DIMENSION);
  o.u32(4);
  o.name("time"); o.u32(1);
  o.name("lev"); o.u32((uint32_t)nz);
  o.name("lat"); o.u32((uint32_t)ny);
  o.name("lon"); o.u32((uint32_t)nx);
  const std::string tag = model_tag ? std::string(model_tag) : std::string();
  o.u32(NC_ATTRIBUTE);
  o.u32(2);
  o.text_att("title", "GEOS-Chem KPP Hnew (KPPHvalue) restart, gckpp_hip");
  o.text_att(kTagAtt, tag);
  o.u32(NC_VARIABLE);
  o.u32(1);
  o.name(kVar);
  o.u32(4);
  for (uint32_t d = 0; d < 4; ++d) o.u32(d);
  o.u32(NC_ATTRIBUTE);
  o.u32(2);
  o.text_att("long_name", "KPP integrator: next internal time step (Hnew)");
  o.text_att("units", "s");
  o.u32(NC_DOUBLE);
  o.u32((uint32_t)(8 * n));
  const size_t begin_at = o.b.size();
  o.u64(0);                                   // begin, set below
  const uint64_t begin = o.b.size();
  for (int k = 0; k < 8; ++k) o.b[begin_at + k] = (unsigned char)(begin >> (8 * (7 - k)));
  for (uint64_t i = 0; i < n; ++i) o.f64(kpph[i]);
  FILE *f = fopen(path, "wb");
  if (!f) return rfail("gckpp_kpph_restart_write: cannot open %s: %s", path, strerror(errno));
  const size_t w = fwrite(o.b.data(), 1, o.b.size(), f);
  if (fclose(f) != 0 || w != o.b.size()) return rfail("gckpp_kpph_restart_write: writing %s failed", path);
  gckpp::set_error("");
  return 0;
}

extern "C" int gckpp_kpph_restart_read(const char *path, int64_t nx, int64_t ny, int64_t nz,
                                       double *kpph, int32_t *found, char *model_tag) {
  if (!path || !kpph || !found) return rfail("gckpp_kpph_restart_read: null argument");
  const int64_t n = nx * ny * nz;
  if (nx <= 0 || ny <= 0 || nz <= 0) return rfail("gckpp_kpph_restart_read: grid sizes out of range");
  *found = 0;
  if (model_tag) memset(model_tag, 0, GCKPP_NAME_LEN);
  std::vector<unsigned char> b;
  {
    FILE *f = fopen(path, "rb");
    if (!f) return rfail("gckpp_kpph_restart_read: cannot open %s: %s", path, strerror(errno));
    unsigned char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + k);
    fclose(f);
  }
  if (b.size() < 8 || b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2))
    return rfail("gckpp_kpph_restart_read: %s is not a netCDF classic file", path);
  const bool off64 = b[3] == 2;
  In r(b);
  r.p = 4;
  r.u32();                                    // numrecs
  // dimensions
  std::vector<uint32_t> dimlen;
  {
    const uint32_t tag = r.u32(), nd = r.u32();
    if (!r.ok || (tag != 0 && tag != NC_DIMENSION)) return rfail("gckpp_kpph_restart_read: %s: bad header", path);
    for (uint32_t i = 0; i < nd && r.ok; ++i) {
      r.name();
      dimlen.push_back(r.u32());
    }
  }
  std::string tag;
  if (!att_list(r, kTagAtt, &tag)) return rfail("gckpp_kpph_restart_read: %s: bad header", path);
  const uint32_t vtag = r.u32(), nv = r.u32();
  if (!r.ok || (vtag != 0 && vtag != NC_VARIABLE)) return rfail("gckpp_kpph_restart_read: %s: bad header", path);
  for (uint32_t i = 0; i < nv; ++i) {
    const std::string nm = r.name();
    const uint32_t nd = r.u32();
    int64_t count = 1;
    for (uint32_t d = 0; d < nd && r.ok; ++d) {
      const uint32_t id = r.u32();
      if (id >= dimlen.size()) return rfail("gckpp_kpph_restart_read: %s: bad dimension id", path);
      count *= dimlen[id] ? (int64_t)dimlen[id] : 1;   // a record dimension: one record
    }
    if (!att_list(r, nullptr, nullptr)) return rfail("gckpp_kpph_restart_read: %s: bad header", path);
    const int t = (int)r.u32();
    r.u32();                                  // vsize
    const uint64_t begin = off64 ? r.u64() : r.u32();
    if (!r.ok) return rfail("gckpp_kpph_restart_read: %s: bad header", path);
    if (nm != kVar) continue;
    if (count != n)
      return rfail("gckpp_kpph_restart_read: %s: Chem_KPPHvalue does not have the grid's size", path);
    if (t != NC_DOUBLE && t != NC_FLOAT)
      return rfail("gckpp_kpph_restart_read: %s: Chem_KPPHvalue is not a floating-point variable", path);
    const size_t ts = (size_t)type_size(t);
    if (begin + ts * (uint64_t)n > b.size()) return rfail("gckpp_kpph_restart_read: %s is truncated", path);
    for (int64_t k = 0; k < n; ++k) {
      const unsigned char *q = b.data() + begin + ts * (uint64_t)k;
      if (t == NC_DOUBLE) {
        uint64_t v = 0;
        for (int j = 0; j < 8; ++j) v = (v << 8) | q[j];
        memcpy(&kpph[k], &v, 8);
      } else {
        uint32_t v = 0;
        for (int j = 0; j < 4; ++j) v = (v << 8) | q[j];
        float x;
        memcpy(&x, &v, 4);
        kpph[k] = x;
      }
    }
    *found = 1;
    if (model_tag) snprintf(model_tag, GCKPP_NAME_LEN, "%s", tag.c_str());
    gckpp::set_error("");
    return 0;
  }
  for (int64_t k = 0; k < n; ++k) kpph[k] = 0.0;   // absent: start from zero (cold)
  gckpp::set_error("");
  return 0;
}
