"""Per-species constants for a mechanism, read from the reference species
database (run/shared/species_database.yml, parsed by
Headers/species_database_mod.F90: KPP_AbsTol :624-632, Background_VV :412-416).

Writes mech/<mech>_species.json with, in KPP species order:
  mw        molecular weight [g/mol]  (MW_g)      -> MW/SR_MW used by rate laws
  henry_k0  Henry_K0 [M/atm], henry_cr Henry_CR [K] -> HENRY_K0/HENRY_CR
  atol      KPP_AbsTol, default 1e-2 (GeosCore/fullchem_mod.F90:2899-2901)
  rtol      KPP_RelTol, default 0.5e-2 (GeosCore/fullchem_mod.F90:2903-2905)
  bg_vv     background VMR: the authored table below where it has the species
            (realistic clean-troposphere values for the species that drive
            the chemistry), else the DB's Background_VV (restart seed values),
            else a default; used only to build synthetic inputs
  bg_src    "authored", "db" or "default"
Run:  python species_db.py <reference-root> <mech>
"""
import json
import re
import sys
from pathlib import Path

import yaml

# Authored background VMRs [mol/mol] for species the DB leaves without one.
# Order-of-magnitude clean/remote-troposphere values; used only to build
# synthetic test inputs.
AUTHORED_BG = {
    "O3": 3.0e-8, "NO": 5.0e-11, "NO2": 2.0e-10, "NO3": 1.0e-12, "N2O5": 1.0e-12,
    "HNO2": 1.0e-11, "HNO3": 3.0e-10, "HNO4": 2.0e-11, "OH": 4.0e-14, "HO2": 1.0e-11,
    "H2O2": 1.0e-9, "CO": 1.0e-7, "CH4": 1.8e-6, "CH2O": 3.0e-10, "MO2": 5.0e-12,
    "MP": 5.0e-10, "PAN": 1.0e-10, "ALD2": 1.0e-10, "ACET": 1.0e-9, "MEK": 1.0e-10,
    "ISOP": 5.0e-10, "MVK": 1.0e-10, "MACR": 5.0e-11, "C2H6": 1.0e-9, "C3H8": 1.0e-10,
    "ALK4": 5.0e-11, "PRPE": 2.0e-11, "SO2": 1.0e-10, "DMS": 5.0e-11, "SO4": 1.0e-10,
    "NH3": 1.0e-10, "HCl": 1.0e-10, "Cl": 1.0e-16, "ClO": 1.0e-13, "HOCl": 1.0e-12,
    "ClNO3": 1.0e-12, "Br": 1.0e-14, "BrO": 1.0e-13, "HBr": 1.0e-12, "HOBr": 1.0e-12,
    "BrNO3": 1.0e-12, "Br2": 1.0e-13, "CHBr3": 1.0e-12, "CH2Br2": 1.0e-12, "CH3Br": 7.0e-12,
    "CH3Cl": 5.0e-10, "N2O": 3.2e-7, "H2O": 1.0e-3, "O": 1.0e-18, "O1D": 1.0e-22,
    "MOH": 1.0e-9, "EOH": 1.0e-10, "GLYX": 1.0e-11, "MGLY": 2.0e-11, "HCOOH": 1.0e-10,
    "ACTA": 2.0e-10, "BENZ": 2.0e-11, "TOLU": 2.0e-11, "XYLE": 1.0e-11, "MTPA": 1.0e-11,
    "LIMO": 1.0e-11, "MTPO": 1.0e-11, "I2": 1.0e-14, "HOI": 1.0e-13, "IO": 1.0e-13,
    "CH3I": 1.0e-12, "OCS": 5.0e-10, "CO2": 4.0e-4, "H2": 5.0e-7, "N2": 0.7808, "O2": 0.2095,
    "Hg0": 1.0e-13, "Hg2": 1.0e-15,
}
DEFAULT_BG = 1.0e-12      # any other stable species
DEFAULT_BG_RADICAL = 1.0e-14  # names ending in O2/OO (peroxy radicals)
DEFAULT_ATOL, DEFAULT_RTOL = 1.0e-2, 0.5e-2
# mechanism-specific defaults for species without KPP_RelTol: Hg uses 1e-2
# (GeosCore/mercury_mod.F90:3842-3843)
DEFAULT_RTOL_BY_MECH = {"fullchem": 0.5e-2, "Hg": 1.0e-2, "carbon": 1.0e-2}


def load_db(refroot):
    path = Path(refroot) / "run" / "shared" / "species_database.yml"
    return yaml.load(path.read_text(), Loader=yaml.SafeLoader)


def species_constants(db, names, mech="fullchem"):
    out = {k: [] for k in ("mw", "henry_k0", "henry_cr", "atol", "rtol", "bg_vv", "bg_src")}
    for n in names:
        e = db.get(n) or {}
        out["mw"].append(float(e.get("MW_g", 0.0) or 0.0))
        out["henry_k0"].append(float(e.get("Henry_K0", 0.0) or 0.0))
        out["henry_cr"].append(float(e.get("Henry_CR", 0.0) or 0.0))
        out["atol"].append(float(e.get("KPP_AbsTol", DEFAULT_ATOL)))
        out["rtol"].append(float(e.get("KPP_RelTol", DEFAULT_RTOL_BY_MECH.get(mech, DEFAULT_RTOL))))
        if n in AUTHORED_BG:
            out["bg_vv"].append(AUTHORED_BG[n])
            out["bg_src"].append("authored")
        elif "Background_VV" in e:
            out["bg_vv"].append(float(e["Background_VV"]))
            out["bg_src"].append("db")
        else:
            v = DEFAULT_BG_RADICAL if re.search(r"(O2|OO\d*)$", n) else DEFAULT_BG
            out["bg_vv"].append(v)
            out["bg_src"].append("default")
    return out


if __name__ == "__main__":
    root, mech = sys.argv[1], sys.argv[2]
    here = Path(__file__).resolve().parent
    m = json.loads((here.parent / "mech" / ("%s.json" % mech)).read_text())
    db = load_db(root)
    sc = species_constants(db, m["species"], mech)
    sc["species"] = m["species"]
    (here.parent / "mech" / ("%s_species.json" % mech)).write_text(json.dumps(sc))
    nz = sum(1 for v in sc["mw"] if v == 0.0)
    print("%s: %d species, %d without MW, %d DB backgrounds, %d non-default ATOL"
          % (mech, len(m["species"]), nz, sc["bg_src"].count("db"),
             sum(1 for a in sc["atol"][:m["nvar"]] if a != DEFAULT_ATOL)))
