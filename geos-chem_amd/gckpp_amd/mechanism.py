"""A KPP mechanism as this project describes it (mech/<name>.json, written by
the mechanism compiler mechgen/eqn_compile.py from the mechanism's KPP
source, plus mech/<name>_species.json from mechgen/species_db.py)."""
import json
import re
from functools import lru_cache
from pathlib import Path

import numpy as np

_MECH_DIR = Path(__file__).resolve().parent.parent / "mech"
MECH_IDS = {"fullchem": 0, "Hg": 1, "carbon": 2}


class Mechanism:
    def __init__(self, name, d=None, species_consts=None):
        """The committed mechanism `name`, or the description d (the JSON
        schema of mechgen/kpp_parse.py and mechgen/eqn_compile.py) with its
        species constants (same schema as mech/<name>_species.json)."""
        if d is None:
            d = json.loads((_MECH_DIR / ("%s.json" % name)).read_text())
        self.name = name
        self.d = d
        self.nspec, self.nvar, self.nfix = d["nspec"], d["nvar"], d["nfix"]
        self.nreact, self.lu_nonzero = d["nreact"], d["lu_nonzero"]
        self.species = d["species"]
        self.reactions = d["reactions"]
        if species_consts is None:
            species_consts = json.loads((_MECH_DIR / ("%s_species.json" % name)).read_text())
        self.species_consts = species_consts
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


def reordered(mech, d):
    """The mechanism of description d, whose species are mech's in another
    order (a compiler's own ordering): species constants permuted by name."""
    pos = {n: i for i, n in enumerate(mech.species)}
    perm = [pos[n] for n in d["species"]]
    sc = {}
    for k, v in mech.species_consts.items():
        sc[k] = [v[i] for i in perm] if isinstance(v, list) and len(v) == mech.nspec else v
    return Mechanism(mech.name, d, sc)
