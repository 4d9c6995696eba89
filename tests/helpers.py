"""Shared test helpers: conversions between the product's designs and the
oracle's problem description, and the error metrics of the parity bar."""
import numpy as np

from oracle import oracle as O
from alpineqms_amd.qms_io import material_slots


def problem(design) -> "O.Problem":
    inp = design.inputs
    mats = np.array([[m.id, *m.mur, *m.sigma] for m in inp.materials], dtype=np.float64).reshape(-1, 7)
    return O.Problem(cells=tuple(inp.cells), h=tuple(inp.h), dt=inp.dt, steps=inp.steps,
                     bc=inp.bc_vector(), traj=np.array(inp.traj, dtype=np.float64),
                     materials=mats, geom=np.asarray(design.geom, dtype=np.float64))


from alpineqms_amd.fields import a_part, rel  # noqa: E402,F401  (shared with bench.py)
from alpineqms_amd.fields import sigma_grad_v as _sgv  # noqa: E402


def sigma_grad_v(p, coef, fields):
    """sigma * grad V on the free A edges (QMS_3D.f90:595-607) for an oracle
    problem and its coefficient array."""
    return _sgv(p.h, coef, fields)
