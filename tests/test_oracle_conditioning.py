"""Conditioning of the solution components (why raw V is not in the 1e-8 bar).

A conductor is tied to the grounded boundary only through the 1e-3 S/m
background, so each conductor's constant potential ("floating" mode) is
determined by conductivities 1e10 times smaller than its own.  Perturbing the
row-scaled right-hand side by a relative 1e-13 moves A, B and sigma*grad(V)
by far less than 1e-8 but moves raw V by orders of magnitude more: V cannot
be pinned to 1e-8 by any FP64 solve, the direct oracle included."""
import numpy as np

from oracle import oracle as O
from alpineqms_amd import configs
from tests.helpers import problem, rel, sigma_grad_v


def test_floating_potential_mode_dominates_v_sensitivity():
    d = configs.small_box_design(14, 21, nbox=3, v=5.0, ramp=True, steps=4)
    p = problem(d)
    coef = O.coeffs(p)
    r = O.run(p, steps=1)
    f = r.fields[-1]
    v = O.step_velocity(p, 2)
    M, b = O.assemble(p, coef, v, p.dt, f)
    S = O.DirectSolver(M)
    x = S.solve(b)
    fx = O.unpack(p, x)
    rng = np.random.default_rng(0)
    bs = b / S.d
    pert = rng.normal(size=b.size)
    pert *= 1e-13 * np.linalg.norm(bs) / np.linalg.norm(pert)
    pert[np.diff(M.indptr) == 1] = 0.0
    f2 = O.unpack(p, S.solve(b + pert * S.d))
    eA = rel(f2[:3], fx[:3])
    eB = rel(O.bfield(p, f2), O.bfield(p, fx))
    eJ = rel(sigma_grad_v(p, coef, f2), sigma_grad_v(p, coef, fx))
    eV = rel(f2[3], fx[3])
    assert eA < 1e-10 and eB < 1e-10 and eJ < 1e-8
    assert eV > 1e3 * max(eA, eB)
    # a 1e-13 right-hand side change already moves raw V past the 1e-8 bar
    assert eV > 1e-8, eV
