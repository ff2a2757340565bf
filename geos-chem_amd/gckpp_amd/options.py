"""Integrator options (ICNTRL/RCNTRL) the GEOS-Chem callers pass to KPP's
Integrate, per mechanism.  These are what a host hands to
gckpp_integrate / gckpp_fullchem_step / gckpp_batch_* (0-based arrays: the
Fortran ICNTRL(k) is element k-1).

Meanings (KPP/fullchem/gckpp_Integrator.F90:220-300):
  ICNTRL(1)  Rosenbrock: 1 autonomous (no dF/dt); carbon: verbose errors
  ICNTRL(2)  0: vector tolerances            ICNTRL(3)  method (4: Rodas3)
  ICNTRL(7)  1: no adjoint                   ICNTRL(15) -1: rates frozen over the step
  ICNTRL(16) carbon only: handling of negative results (0 keep, 1 zero, 2-3 error)
  RCNTRL(3)  Hstart (per cell in this library: the hstart argument)
"""
import numpy as np


def fullchem_icntrl():
    """fullchem_AR_SetIntegratorOptions with autoreduce off
    (KPP/fullchem/fullchem_AutoReduceFuncs.F90:249-287): ICNTRL(1)=1,
    ICNTRL(3)=4, ICNTRL(7)=1, ICNTRL(15)=-1; RCNTRL zero (RCNTRL(3), the
    previous step's KPPHvalue, is passed per cell)."""
    ic = np.zeros(20, dtype=np.int32)
    ic[0], ic[2], ic[6], ic[14] = 1, 4, 1, -1
    return ic, np.zeros(20)


def hg_icntrl():
    """The Hg chemistry call (GeosCore/mercury_mod.F90:951-969): autonomous,
    vector tolerances, Rodas3, no adjoint, rates frozen."""
    ic = np.zeros(20, dtype=np.int32)
    ic[0], ic[2], ic[6], ic[14] = 1, 4, 1, -1
    return ic, np.zeros(20)


def carbon_icntrl():
    """The carbon chemistry call (GeosCore/carbon_gases_mod.F90:453-456):
    verbose errors, negatives kept (ICNTRL(16)=0), rates frozen."""
    ic = np.zeros(20, dtype=np.int32)
    ic[0], ic[14] = 1, -1
    return ic, np.zeros(20)


def icntrl_for(mech_name):
    """(ICNTRL, RCNTRL) of the caller of mechanism mech_name."""
    return {"fullchem": fullchem_icntrl, "Hg": hg_icntrl, "carbon": carbon_icntrl}[mech_name]()
