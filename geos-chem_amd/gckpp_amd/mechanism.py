"""A KPP mechanism as this project describes it (mech/<name>.json, written by
mechgen/kpp_parse.py from the reference's KPP tables, plus
mech/<name>_species.json from mechgen/species_db.py)."""
import json
import re
from functools import lru_cache
from pathlib import Path

import numpy as np

_MECH_DIR = Path(__file__).resolve().parent.parent / "mech"
MECH_IDS = {"fullchem": 0, "Hg": 1, "carbon": 2}


class Mechanism:
    def __init__(self, name):
        d = json.loads((_MECH_DIR / ("%s.json" % name)).read_text())
        self.name = name
        self.d = d
        self.nspec, self.nvar, self.nfix = d["nspec"], d["nvar"], d["nfix"]
        self.nreact, self.lu_nonzero = d["nreact"], d["lu_nonzero"]
        self.species = d["species"]
        self.reactions = d["reactions"]
        self.species_consts = json.loads((_MECH_DIR / ("%s_species.json" % name)).read_text())
        self.lu_irow = np.asarray(d["lu_irow"], dtype=np.int32)
        self.lu_icol = np.asarray(d["lu_icol"], dtype=np.int32)
        self.lu_crow = np.asarray(d["lu_crow"], dtype=np.int32)
        self.lu_diag = np.asarray(d["lu_diag"], dtype=np.int32)
        self.index = {n: i for i, n in enumerate(self.species)}

    @property
    def mech_id(self):
        return MECH_IDS[self.name]

    def atol(self):
        return np.asarray(self.species_consts["atol"][: self.nvar], dtype=np.float64)

    def rtol(self):
        return np.asarray(self.species_consts["rtol"][: self.nvar], dtype=np.float64)

    def photol_slots(self):
        """{0-based PHOTOL slot: name of the first photolysed species}"""
        out = {}
        for rx in self.reactions:
            for m in re.finditer(r"PHOTOL\[(\d+)\]", rx["rate"]):
                s = int(m.group(1))
                if s not in out and rx["reactants"]:
                    out[s] = self.species[rx["reactants"][0][0]]
        return out

    def inlined_rates(self):
        """{reaction: value} for reactions whose constant rate KPP substituted
        directly into Fun/Jac (RCONST(r) is never set for them)."""
        return {r: float(rx["rate"]) for r, rx in enumerate(self.reactions) if rx["inlined"]}

    def family_species(self):
        """Indices of prod/loss dummy species (ATOL 1e25 in the species DB)."""
        return [i for i in range(self.nvar) if self.species_consts["atol"][i] >= 1e20]


@lru_cache(maxsize=None)
def load(name):
    return Mechanism(name)
