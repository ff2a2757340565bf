"""Inputs that drive Integrate into its step-size error branches, shared by
the CPU pin of the C oracle and the GPU test.

Step too small (IERR -7, KPP/fullchem/gckpp_Integrator.F90:659-662): the
time loop returns -7 when T + 0.1*H == T at the start of a step.
Integrating from TIN = 1e12 s with the cold start H = 1e-5 s hits it on the
first check (the spacing of doubles at 1e12 is 1.2e-4 s).  From TIN = 2**42
s (spacing 9.8e-4 s) with Hstart = 1 s the first check passes; cells whose
controller then accepts steps shorter than about 4.9e-3 s stop with -7 a few
steps in, the others finish.

Singular matrix (ros_PrepareMatrix, :1921-1999; KppDecomp's pivot check,
gckpp_LinearAlgebra.F90:61-65): Ghimj = -Jac0 + 1/(H*gamma1) on the
diagonal.  CH2I2 is consumed only by its photolysis, so its Jacobian
diagonal is minus that rate.  A photolysis rate of -2048 /s (a rate the
reference accepts as given) with Hstart = 2**-10 s makes the diagonal
-2048 + 1/(2**-10 * 0.5) = 0 exactly: the decomposition reports row 1
singular, H is halved (ISTATUS(8) = 1) and the step proceeds.  A -8 needs
six consecutive singular decompositions at H, H/2, ..., H/32; with a fixed
Jacobian the diagonal can vanish at one of those H only, so -8 is not
reachable from any input and is covered by reading the code alone.
"""
import numpy as np

T_TINY_FIRST = 1.0e12
T_TINY_LATER = float(2 ** 42)
H_LATER = 1.0
H_SING = 2.0 ** -10


def singular_cells(m, w, rct):
    """(conc, rct, hstart, tout) of cells where the first decomposition is
    singular (even cells) or not (odd cells: the rate off by one part in
    2**40)."""
    assert m.name == "fullchem"
    k = m.index["CH2I2"]
    rx = [r for r, x in enumerate(m.reactions) if any(s == k for s, _ in x["reactants"])]
    assert len(rx) == 1 and "PHOTOL" in m.reactions[rx[0]]["rate"]
    r = rx[0]
    n = w.ncell
    rct = np.array(rct, dtype=np.float64, copy=True)
    rate = np.where(np.arange(n) % 2 == 0, -2048.0, -2048.0 * (1.0 + 2.0 ** -40))
    rct[r] = rate
    conc = np.array(w.conc, dtype=np.float64, copy=True)
    hstart = np.full(n, H_SING)
    return conc, rct, hstart, 4 * H_SING
