"""Synthetic chemistry-step inputs (the workloads the benchmarks and parity
tests run), built in the cell-fastest layout of the C ABI.

A Workload holds, for ncell cells:
  met    [NMET][ncell]      TEMP, PRESS, NUMDEN, H2O, SUNCOS, ... (cell_fields.h)
  photol [200][ncell]       PHOTOL(1:200) [1/s]
  het    [NHET][1]          one clear-sky, aerosol-free HetState for all cells
  conc   [NSPEC][ncell]     KPP concentrations [molec/cm3]
  hstart [ncell]            RCNTRL(3) per cell (0 -> Hstart = 1e-5 s)

Box cells (``box_cells``): T ~ U(190,310) K, P ~ logU(1,1013) hPa, NUMDEN from
the ideal gas law (BOLTZ, Headers/physconstants.F90:67), H2O VMR ~
logU(1e-6,3e-2); the first half dark (J=0, SUNCOS=-0.5), the second half lit
with J_i = J0_i*cosSZA^0.8, cosSZA ~ U(0,1]; every species log-uniform within
+/-1 decade of its background VMR.

Grid cells (``grid_cells``): an NX x NY x NZ lat/lon grid, T from a lapse-rate
profile plus N(0,5 K), pressure from the 72-level hybrid grid
(GeosUtil/pressure_mod.F90:680-730, centres as in :312-320), cosSZA from
lat/lon at 12:00 UTC on day 172, log-normal smooth concentration fields
(sigma 0.5 decade, correlation length 5 cells).
"""
import json
import math
from pathlib import Path

import numpy as np

from . import fields
from .mechanism import Mechanism

BOLTZ = 1.38064852e-23      # Headers/physconstants.F90:67
_HERE = Path(__file__).resolve().parent

# Authored overhead-sun photolysis frequencies [1/s], by photolysed species.
_J0_BY_SPECIES = {
    "O2": 1.0e-11, "NO2": 9.0e-3, "NO3": 1.0e-1, "HNO2": 1.8e-3, "NO": 1.0e-8,
    "H2O2": 8.0e-6, "MP": 6.0e-6, "CH2O": 3.5e-5, "HNO3": 7.0e-7, "HNO4": 5.0e-6,
    "N2O5": 3.0e-5, "Cl2": 2.5e-3, "Br2": 3.5e-2, "BrO": 4.0e-2, "BrCl": 1.2e-2,
    "HOBr": 2.0e-3, "HOCl": 2.5e-4, "ClO": 1.0e-5, "OClO": 7.0e-2, "Cl2O2": 1.5e-3,
    "ClNO2": 3.0e-4, "ClNO3": 4.0e-5, "BrNO3": 1.2e-3, "BrNO2": 2.0e-3,
    "I2": 1.2e-1, "HOI": 5.0e-3, "IO": 1.0e-1, "OIO": 1.0e-1, "IONO": 2.0e-2,
    "IONO2": 1.0e-2, "GLYX": 8.0e-5, "MGLY": 1.2e-4, "ALD2": 4.0e-6, "ACET": 5.0e-7,
    "PAN": 3.0e-7, "MVK": 1.5e-6, "MACR": 2.0e-6,
}
_J0_O3 = {1: 4.5e-4, 2: 3.0e-5}   # PHOTOL(2): O3 -> O, PHOTOL(3): O3 -> O1D
_J0_DEFAULT = 1.0e-5


def j0_table(mech):
    """J0[200] for the PHOTOL slots the mechanism uses (0 for unused slots)."""
    j0 = np.zeros(fields.NPHOTOL)
    for slot, spc in mech.photol_slots().items():
        if mech.name == "fullchem" and slot in _J0_O3:
            j0[slot] = _J0_O3[slot]
        else:
            j0[slot] = _J0_BY_SPECIES.get(spc, _J0_DEFAULT)
    return j0


def clear_sky_het(ncell_or_one=1):
    """One HetState record: no cloud, no aerosol (areas, radii, volumes and
    water zero), clear-sky fraction 1, no PSC/NAT, het rates on."""
    h = np.zeros((fields.NHET, ncell_or_one))
    F = fields.HET
    h[F["NAEROTYPE"][0]] = 14
    h[F["clearFr"][0]] = 1.0
    h[F["fupdateHOBr"][0]] = 1.0
    h[F["fupdateHOCl"][0]] = 1.0
    h[F["gamma_HO2"][0]] = 0.2
    h[F["pHCloud"][0]] = 4.5
    h[F["pHSSA"][0]:F["pHSSA"][0] + 2] = 5.0
    h[F["H_plus"][0]] = 1.0e-5
    h[F["H_conc_Sul"][0]] = 1.0e-5
    h[F["H_conc_LCl"][0]] = 1.0e-5
    h[F["H_conc_ICl"][0]] = 1.0e-5
    h[F["H_conc_SSA"][0]] = 1.0e-5
    h[F["H_conc_SSC"][0]] = 1.0e-5
    h[F["OMOC_POA"][0]] = 1.4
    h[F["OMOC_OPOA"][0]] = 2.1
    return h


class Workload:
    def __init__(self, mech, met, photol, het, conc, hstart, dt=1200.0, name=""):
        self.mech, self.met, self.photol, self.het = mech, met, photol, het
        self.conc, self.hstart, self.dt, self.name = conc, hstart, dt, name

    @property
    def ncell(self):
        return self.conc.shape[1]

    def subset(self, idx):
        idx = np.asarray(idx)
        het = self.het if self.het.shape[1] == 1 else self.het[:, idx]
        return Workload(self.mech, np.ascontiguousarray(self.met[:, idx]),
                        np.ascontiguousarray(self.photol[:, idx]), het,
                        np.ascontiguousarray(self.conc[:, idx]),
                        np.ascontiguousarray(self.hstart[idx]), self.dt, self.name)


def _fill_met(met, temp, press, numden, h2o, suncos):
    M = fields.MET
    met[M["TEMP"][0]] = temp
    met[M["PRESS"][0]] = press
    met[M["NUMDEN"][0]] = numden
    met[M["H2O"][0]] = h2o
    met[M["SUNCOS"][0]] = suncos
    met[M["TROP"][0]] = (press > 200.0).astype(float)


def _species_conc(mech, numden, logdev):
    """conc[NSPEC][ncell]: bg_i * NUMDEN * 10**logdev[i], fixed species at
    their fixed mixing ratios, prod/loss dummy species (ATOL 1e25) zero."""
    sc = mech.species_consts
    bg = np.asarray(sc["bg_vv"])
    conc = bg[:, None] * numden[None, :] * 10.0 ** logdev
    for i, name in enumerate(mech.species):
        if i >= mech.nvar:
            conc[i] = bg[i] * numden      # fixed species: no perturbation
        elif sc["atol"][i] >= 1.0e20:
            conc[i] = 0.0                 # prod/loss family dummies
    return conc


def _carbon_inputs(mech, met, conc, rng):
    """The carbon mechanism's per-cell rate inputs, as
    carbon_ComputeRateConstants sets them (KPP/carbon/carbon_Funcs.F90:171-245):
    in tropospheric cells (TROP=1) fixed OH and Cl, k_Trop(1) = P(CO) from CH4,
    k_Trop(2) = k(CO+OH), k_Trop(3) = P(CO) from NMVOC; in stratospheric
    cells k_Strat(1), k_Strat(3) first-order CH4 and CO losses and k_Strat(2)
    = P(CO) there; the dummy species are 1 (:102-109)."""
    M = fields.MET
    n = conc.shape[1]
    trop = met[M["TROP"][0]] > 0.0
    lu = lambda lo, hi: 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), n)
    idx = {s: i for i, s in enumerate(mech.species)}
    for d in ("DummyCH4trop", "DummyCH4strat", "DummyNMVOC"):
        conc[idx[d]] = 1.0
    conc[idx["FixedOH"]] = np.where(trop, lu(1e5, 5e6), 1.0)
    conc[idx["FixedCl"]] = np.where(trop, lu(1e2, 1e4), 1.0)
    kt, ks = M["K_TROP"][0], M["K_STRAT"][0]
    met[kt:kt + 5] = 0.0
    met[ks:ks + 5] = 0.0
    met[kt + 0] = np.where(trop, lu(1e1, 1e4), 0.0)
    met[kt + 1] = np.where(trop, lu(1.5e-13, 3e-13), 0.0)
    met[kt + 2] = np.where(trop, lu(1e1, 1e4), 0.0)
    met[ks + 0] = np.where(trop, 0.0, lu(1e-10, 1e-7))
    met[ks + 1] = np.where(trop, 0.0, lu(1e0, 1e3))
    met[ks + 2] = np.where(trop, 0.0, lu(1e-10, 1e-7))


def box_cells(mech, ncell=4096, seed=0, dt=1200.0):
    rng = np.random.default_rng(seed)
    temp = rng.uniform(190.0, 310.0, ncell)
    press = 10.0 ** rng.uniform(0.0, math.log10(1013.0), ncell)
    numden = press * 100.0 / (BOLTZ * temp) * 1.0e-6
    h2o = 10.0 ** rng.uniform(-6.0, math.log10(3.0e-2), ncell) * numden
    dark = np.arange(ncell) < ncell // 2
    cossza = 1.0 - rng.uniform(0.0, 1.0, ncell)          # (0, 1]
    suncos = np.where(dark, -0.5, cossza)
    met = np.zeros((fields.NMET, ncell))
    _fill_met(met, temp, press, numden, h2o, suncos)
    j0 = j0_table(mech)
    photol = np.where(dark[None, :], 0.0, j0[:, None] * np.clip(cossza, 0, 1)[None, :] ** 0.8)
    logdev = rng.uniform(-1.0, 1.0, (mech.nspec, ncell))
    conc = _species_conc(mech, numden, logdev)
    if mech.name == "carbon":
        _carbon_inputs(mech, met, conc, rng)
    return Workload(mech, met, np.ascontiguousarray(photol), clear_sky_het(),
                    conc, np.zeros(ncell), dt, "box%d_seed%d" % (ncell, seed))


def _smooth_field(rng, shape, corr_cells):
    """Unit-variance Gaussian random field with a Gaussian correlation of
    ``corr_cells`` grid cells, by FFT filtering of white noise."""
    noise = rng.standard_normal(shape)
    spec = np.fft.rfftn(noise)
    ks = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    grids = np.meshgrid(*ks, indexing="ij")
    k2 = sum(g * g for g in grids)
    filt = np.exp(-0.5 * k2 * (2 * math.pi * corr_cells) ** 2 / 2.0)
    f = np.fft.irfftn(spec * filt, s=shape)
    f -= f.mean()
    return f / (f.std() + 1e-300)


def grid_cells(mech, nx=72, ny=46, nz=72, seed=1, dt=1200.0, nspec_fields=None, hour_utc=12.0):
    """Cells of an nx*ny*nz lat/lon grid, ordered I fastest then J then L;
    the sun as at hour_utc on day 172."""
    rng = np.random.default_rng(seed)
    g = json.loads((_HERE.parent / "mech" / "grid72.json").read_text())
    ap, bp = np.asarray(g["ap_hpa"]), np.asarray(g["bp"])
    nz = min(nz, 72)
    lon = -180.0 + (np.arange(nx) + 0.5) * 360.0 / nx
    lat = -90.0 + (np.arange(ny) + 0.5) * 180.0 / ny
    psurf = 1013.25 - 0.0 * lat[:, None] - 0.0 * lon[None, :]      # [ny, nx]
    pe = ap[:, None, None] + bp[:, None, None] * psurf[None]          # [73, ny, nx]
    pc = 0.5 * (pe[:-1] + pe[1:])[:nz]                              # [nz, ny, nx]
    z = -7.6 * np.log(pc / 1013.25)                                  # km, scale height 7.6
    tsurf = 300.0 - 40.0 * np.sin(np.radians(lat)) ** 2              # [ny]
    tprof = np.maximum(tsurf[None, :, None] - 6.5 * z, 210.0)
    tprof = np.where(z > 20.0, tprof + 2.0 * (z - 20.0), tprof)
    temp = tprof + rng.normal(0.0, 5.0, tprof.shape)
    temp = np.clip(temp, 180.0, 320.0)
    numden = pc * 100.0 / (BOLTZ * temp) * 1.0e-6
    h2o_vmr = np.clip(1.0e-2 * np.exp(-z / 2.5), 4.0e-6, 3.0e-2)
    h2o = h2o_vmr * numden
    # solar zenith angle at 12:00 UTC, day 172 (solstice declination)
    decl = math.radians(23.44) * math.sin(2 * math.pi * (284 + 172) / 365.0)
    hour_angle = np.radians(lon + 15.0 * (hour_utc - 12.0))   # local solar hour angle
    latr = np.radians(lat)
    cossza2d = (np.sin(latr)[:, None] * math.sin(decl)
                + np.cos(latr)[:, None] * math.cos(decl) * np.cos(hour_angle)[None, :])
    cossza = np.broadcast_to(cossza2d[None], (nz, ny, nx))
    ncell = nx * ny * nz
    met = np.zeros((fields.NMET, ncell))
    flat = lambda a: np.ascontiguousarray(a.reshape(-1))
    _fill_met(met, flat(temp), flat(pc), flat(numden), flat(h2o), flat(cossza))
    j0 = j0_table(mech)
    lit = flat(cossza) > 0.0
    photol = np.where(lit[None, :], j0[:, None] * np.clip(flat(cossza), 0, 1)[None, :] ** 0.8, 0.0)
    # log-normal smooth fields: a few shared fields mixed per species keeps
    # this cheap for 356 species on a million cells
    nbase = nspec_fields or 8
    base = np.stack([flat(_smooth_field(rng, (nz, ny, nx), 5.0)) for _ in range(nbase)])
    mix = rng.standard_normal((mech.nspec, nbase))
    mix /= np.linalg.norm(mix, axis=1, keepdims=True)
    logdev = 0.5 * (mix @ base)
    conc = _species_conc(mech, flat(numden), logdev)
    if mech.name == "carbon":
        _carbon_inputs(mech, met, conc, rng)
    return Workload(mech, met, np.ascontiguousarray(photol), clear_sky_het(),
                    conc, np.zeros(ncell), dt,
                    "grid%dx%dx%d_seed%d%s" % (nx, ny, nz, seed,
                                               "" if hour_utc == 12.0 else "_%gUTC" % hour_utc))


# Emission rates [molec/cm3/s] of the boundary-layer cells (P > 700 hPa),
# log-uniform between the bounds: a synthetic stand-in for the emission and
# transport operators that run between two chemistry steps.
_EMIS = {"NO": (1e4, 3e6), "CO": (1e5, 3e7), "ISOP": (1e3, 3e6), "CH2O": (1e3, 1e6),
         "SO2": (1e3, 1e6), "ALD2": (1e3, 3e5)}


def perturb(mech, conc, met, dt=1200.0, seed=0, mixing=0.05):
    """The concentrations after the operators between two chemistry steps:
    every variable species scaled by exp(N(0, mixing)) (transport), then the
    emissions of _EMIS over dt added in the boundary layer.  Returns a new
    array."""
    rng = np.random.default_rng(seed)
    n = conc.shape[1]
    out = np.array(conc, copy=True)
    nv = mech.nvar
    out[:nv] *= np.exp(rng.normal(0.0, mixing, (nv, n)))
    bl = met[fields.MET["PRESS"][0]] > 700.0
    for name, (lo, hi) in _EMIS.items():
        if name in mech.species:
            e = 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), n)
            out[mech.species.index(name)] += np.where(bl, e * dt, 0.0)
    return out


def perturb_dev(mech, conc, met, gen, dt=1200.0, mixing=0.05):
    """perturb() on device tensors, in place: conc [NSPEC][n] and met
    [NMET][n] torch tensors on the GPU, gen a torch.Generator on the same
    device (the same distributions as perturb, other random draws)."""
    import torch
    nv = mech.nvar
    n = conc.shape[1]
    conc[:nv] *= torch.exp(torch.randn((nv, n), generator=gen, device=conc.device,
                                       dtype=conc.dtype) * mixing)
    bl = met[fields.MET["PRESS"][0]] > 700.0
    for name, (lo, hi) in _EMIS.items():
        if name in mech.species:
            u = torch.rand(n, generator=gen, device=conc.device, dtype=conc.dtype)
            e = 10.0 ** (np.log10(lo) + u * (np.log10(hi) - np.log10(lo)))
            conc[mech.species.index(name)] += torch.where(bl, e * dt, torch.zeros_like(e))
    return conc


def mixed_grid(mechs, nx=144, ny=91, nz=72, seed=3, split=(0.7, 0.2, 0.1), hour_utc=12.0):
    """Cells of one nx*ny*nz grid assigned at random to the mechanisms mechs
    (e.g. fullchem, Hg, carbon) with probabilities split; returns one
    Workload per mechanism (its cells, in grid order) and the assignment."""
    rng = np.random.default_rng(seed + 7919)
    n = nx * ny * nz
    assign = rng.choice(len(mechs), size=n, p=np.asarray(split) / np.sum(split))
    parts = []
    for k, m in enumerate(mechs):
        idx = np.nonzero(assign == k)[0]
        parts.append(grid_cells(m, nx, ny, nz, seed=seed, hour_utc=hour_utc).subset(idx))
    return parts, assign


def rich_het(rng, ncell):
    """Per-cell HetState records with cloud, ice and every aerosol type
    present, random pH, halide concentrations and flags: exercises every
    branch of the heterogeneous rate laws (used by the rate parity tests)."""
    F = fields.HET
    h = np.zeros((fields.NHET, ncell))
    u = lambda lo, hi: rng.uniform(lo, hi, ncell)
    lu = lambda lo, hi: 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), ncell)
    flag = lambda p: (rng.uniform(0, 1, ncell) < p).astype(float)
    for name in ("natSurface", "pscBox", "SSA_is_Alk", "SSC_is_Alk", "cloudBox", "TurnOffHetRates"):
        h[F[name][0]] = flag(0.3)
    h[F["stratBox"][0]] = flag(0.3)
    acid = flag(0.6)
    h[F["SSA_is_Acid"][0]] = acid
    h[F["SSC_is_Acid"][0]] = flag(0.6)
    h[F["NAEROTYPE"][0]] = 14
    for t in range(14):
        h[F["xArea"][0] + t] = lu(1e-9, 1e-5)
        h[F["xRadi"][0] + t] = lu(1e-6, 1e-3)
        h[F["xVol"][0] + t] = lu(1e-14, 1e-10)
        h[F["xH2O"][0] + t] = lu(1e-15, 1e-11)
        h[F["wetArea"][0] + t] = lu(1e-9, 1e-5)
    h[F["aClArea"][0]] = lu(1e-9, 1e-5)
    h[F["aClRadi"][0]] = lu(1e-6, 1e-4)
    h[F["aClVol"][0]] = lu(1e-14, 1e-10)
    cld = u(0.0, 1.0)
    h[F["cldFr"][0]] = cld
    h[F["clearFr"][0]] = 1.0 - cld
    h[F["aLiq"][0]] = lu(1e-7, 1e-4)
    h[F["aIce"][0]] = lu(1e-8, 1e-5)
    h[F["rLiq"][0]] = lu(5e-4, 2e-3)
    h[F["rIce"][0]] = lu(1e-3, 5e-3)
    h[F["aWater"][0]:F["aWater"][0] + 2] = lu(1e-12, 1e-8)
    for name in ("Br_conc_CldA", "Br_conc_CldC", "Br_conc_CldG", "Br_conc_Cld", "Br_conc_SSA",
                 "Br_conc_SSC", "Cl_conc_CldA", "Cl_conc_CldC", "Cl_conc_CldG", "Cl_conc_Cld",
                 "Cl_conc_SSA", "Cl_conc_SSC", "NIT_conc_SSA", "NIT_conc_SSC"):
        h[F[name][0]] = lu(1e-8, 1e-1)
    for name in ("Br_over_Cl_Cld", "Br_over_Cl_SSA", "Br_over_Cl_SSC"):
        h[F[name][0]] = lu(1e-5, 1e-2)
    for name in ("frac_Br_CldA", "frac_Br_CldC", "frac_Br_CldG", "frac_Cl_CldA", "frac_Cl_CldC",
                 "frac_Cl_CldG", "frac_HSO3_aq", "frac_SALACL", "frac_SO3_aq", "f_Acid_SSA",
                 "f_Acid_SSC", "f_Alk_SSA", "f_Alk_SSC", "HBr_theta", "HCl_theta", "HNO3_theta"):
        h[F[name][0]] = u(0.0, 1.0)
    h[F["fupdateHOBr"][0]] = 1.0
    h[F["fupdateHOCl"][0]] = 1.0
    h[F["gamma_HO2"][0]] = u(0.0, 0.3)
    for name in ("H_conc_ICl", "H_conc_LCl", "H_conc_SSA", "H_conc_SSC", "H_conc_Sul", "H_plus"):
        h[F[name][0]] = lu(1e-9, 1e-1)
    for name in ("HSO3_aq", "SO3_aq", "TSO3_aq", "HSO4_molal", "NO3_molal", "SO4_molal"):
        h[F[name][0]] = lu(1e-9, 1e-2)
    h[F["HSO3m"][0]] = u(0.0, 1.0)
    h[F["SO3mm"][0]] = u(0.0, 1.0)
    h[F["KHETI_SLA"][0]:F["KHETI_SLA"][0] + 11] = lu(1e-6, 1e-2)
    h[F["pHCloud"][0]] = u(1.0, 7.0)
    h[F["pHSSA"][0]:F["pHSSA"][0] + 2] = u(1.0, 8.0)
    h[F["OMOC_POA"][0]] = u(1.2, 2.0)
    h[F["OMOC_OPOA"][0]] = u(1.6, 2.4)
    h[F["qIce"][0]] = lu(1e-7, 1e-4)
    h[F["qLiq"][0]] = lu(1e-7, 1e-3)
    h[F["vAir"][0]] = lu(1e12, 1e15)
    h[F["vIce"][0]] = lu(1e-9, 1e-6)
    h[F["vLiq"][0]] = lu(1e-8, 1e-5)
    h[F["fracOrgAer"][0]] = u(0.0, 1.0)
    h[F["fracInorgAer"][0]] = u(0.0, 1.0)
    return h


def model_hetin(rng, ncell):
    """The model's per-cell fields that fullchem_SetStateHet reads
    (hetin[NHIN][ncell], fields.HIN): clouds (none, liquid, ice, mixed),
    temperatures across the ice-radius regimes (202 K, 217 K), land and
    ocean boxes, aerosol of every type, cloud and aerosol pH (some <= 0),
    sea-salt alkalinity with empty boxes and ratios above one, PSC states
    in and out of the stratosphere."""
    F = fields.HIN
    h = np.zeros((fields.NHIN, ncell))
    u = lambda lo, hi: rng.uniform(lo, hi, ncell)
    lu = lambda lo, hi: 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), ncell)
    some = lambda p: rng.uniform(0, 1, ncell) < p
    put = lambda name, v, k=0: h.__setitem__(F[name][0] + k, v)
    cldf = np.where(some(0.3), 0.0, u(-0.02, 1.02))
    put("CLDF", cldf)
    put("QI", np.where(some(0.4), 0.0, lu(1e-8, 1e-4)))
    put("QL", np.where(some(0.3), 0.0, lu(1e-8, 1e-3)))
    airvol = lu(1e8, 1e12)
    airden = lu(0.01, 1.3)
    put("AIRVOL", airvol)
    put("AIRDEN", airden)
    put("AD", airden * airvol)
    land = u(0, 1)
    put("FRLAND", land)
    put("FROCEAN", (1.0 - land) * u(0, 1))
    put("T", u(180.0, 310.0))
    put("aClArea", np.where(some(0.1), 0.0, lu(1e-9, 1e-5)))
    put("aClRadi", lu(1e-6, 1e-4))
    for k in range(2):
        put("IsorropAeroH2O", lu(1e-12, 1e-8), k)
        put("IsorropAeropH", u(-1.0, 8.0), k)
    for k in range(14):
        put("AeroArea", np.where(some(0.05), 0.0, lu(1e-9, 1e-5)), k)
        put("AeroRadi", lu(1e-6, 1e-3), k)
        put("WetAeroArea", lu(1e-9, 1e-5), k)
        put("AeroH2O", lu(1e-10, 1e-4), k)
    # no sea-salt water without sea-salt area (the NO3 uptake on sea salt
    # divides the water by the aerosol volume)
    h[F["IsorropAeroH2O"][0]][h[F["aClArea"][0]] == 0] = 0.0
    h[F["IsorropAeroH2O"][0] + 1][h[F["AeroArea"][0] + 11] == 0] = 0.0
    put("OMOC_POA", u(1.4, 2.1))
    put("OMOC_OPOA", u(1.4, 2.1))
    put("HSO3_aq", np.where(some(0.3), 0.0, lu(1e-9, 1e-4)))
    put("SO3_aq", np.where(some(0.3), 0.0, lu(1e-12, 1e-6)))
    put("IsorropBisulfate", u(0.0, 5.0))
    put("IsorropNitrate", u(0.0, 5.0))
    put("IsorropSulfate", u(0.0, 5.0))
    put("IsorropHplus", lu(1e-8, 1e-1))
    put("pHCloud", np.where(some(0.1), u(-1.0, 0.0), u(2.0, 7.5)))
    sala, salc = lu(1e-12, 1e-8), lu(1e-12, 1e-8)
    put("SALA", np.where(some(0.1), 0.0, sala))
    put("SALAAL", sala * u(0.0, 1.2))
    put("SALC", np.where(some(0.1), 0.0, salc))
    put("SALCAL", salc * u(0.0, 1.2))
    put("fupdateHOBr", u(0.0, 1.0))
    put("fupdateHOCl", u(0.0, 1.0))
    for k in range(11):
        put("KHETI_SLA", u(0.0, 0.3), k)
    put("InStratosphere", some(0.3).astype(float))
    put("STATE_PSC", rng.integers(0, 4, ncell).astype(float))
    return h
