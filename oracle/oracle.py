"""CPU oracle for the QMS_3D time-stepping core.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, never by the product path.

This module drives the C restatement in ``qms_oracle.c`` (built to
``oracle/libqms_oracle.so`` by ``oracle/Makefile``) and replaces the
reference's MKL PARDISO call with an FP64 sparse direct solve
(SciPy SuperLU, the closest algorithmic analogue available here).  It runs the
whole program flow of ``3D/QMS_3D.f90``: steady solve (``:424-1415``), then
``time_steps`` transient solves (``:1424-2302``), each followed by the unpack
(``:2211-2234``) and the cell-centred H (``:2237-2267``).

Array conventions: node-sized fields are numpy arrays of shape (Nz, Ny, Nx)
(x fastest in memory, the Fortran layout of the reference); H arrays have
shape (cz, cy, cx).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int32)


def build() -> str:
    """Compile the C restatement (gcc) and return the library path."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libqms_oracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libqms_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.qo_ndims.restype = ctypes.c_int64
        L.qo_ndims.argtypes = [ctypes.c_int] * 3
        L.qo_nnz.restype = ctypes.c_int64
        L.qo_nnz.argtypes = [ctypes.c_int] * 3
        L.qo_exact_nn.restype = ctypes.c_int64
        L.qo_exact_nn.argtypes = [ctypes.c_int] * 3
        L.qo_coeffs.restype = ctypes.c_int
        L.qo_coeffs.argtypes = [ctypes.c_int] * 3 + [_d, ctypes.c_int, _d, _d]
        L.qo_assemble.restype = ctypes.c_int64
        L.qo_assemble.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 3
                                  + [_d, _d, _d, ctypes.c_double, _d, _i64, _i32, _d, _d])
        L.qo_bc_fields.restype = None
        L.qo_bc_fields.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [_d, _d, _d]
        L.qo_unpack.restype = None
        L.qo_unpack.argtypes = [ctypes.c_int] * 3 + [_d, _d]
        L.qo_pack.restype = None
        L.qo_pack.argtypes = [ctypes.c_int] * 3 + [_d, _d]
        L.qo_hfield.restype = None
        L.qo_hfield.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [_d, _d, _d]
        L.qo_trajectory.restype = ctypes.c_double
        L.qo_trajectory.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.qo_step_velocity.restype = ctypes.c_double
        L.qo_step_velocity.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        _LIB = L
    return _LIB


def _p(a: np.ndarray, t=_d):
    return a.ctypes.data_as(t)


@dataclass
class Problem:
    """Everything the reference reads from inputs.txt and geom.bin
    (``QMS_3D.f90:121-165``, ``:229-232``)."""
    cells: tuple                      # (cx, cy, cz)
    h: tuple                          # (dx, dy, dz)
    dt: float
    steps: int
    bc: np.ndarray                    # (12,) B0 then G rows (gx*, gy*, gz*)
    traj: np.ndarray                  # (3, 3): per axis (v0, a, n_a)
    materials: np.ndarray             # (nmat, 7): id, mur xyz, sigma xyz
    geom: np.ndarray                  # (cz, cy, cx) float64 material IDs

    @property
    def nodes(self):
        return tuple(c + 1 for c in self.cells)


def ndims(cells) -> int:
    return int(lib().qo_ndims(*cells))


def nnz(cells) -> int:
    return int(lib().qo_nnz(*cells))


def exact_nn(cells) -> int:
    return int(lib().qo_exact_nn(*cells))


def coeffs(p: Problem) -> np.ndarray:
    """(6, Nz, Ny, Nx): Mu_x, Mu_y, Mu_z, Sigma_x, Sigma_y, Sigma_z
    (``QMS_3D.f90:211-269``)."""
    cx, cy, cz = p.cells
    Nx, Ny, Nz = p.nodes
    geom = np.ascontiguousarray(p.geom, dtype=np.float64)
    mat = np.ascontiguousarray(p.materials, dtype=np.float64).reshape(-1, 7)
    out = np.empty((6, Nz, Ny, Nx), dtype=np.float64)
    rc = lib().qo_coeffs(cx, cy, cz, _p(geom), mat.shape[0], _p(mat), _p(out))
    if rc != 0:
        raise ValueError("geometry holds a material ID outside the table")
    return out


def assemble(p: Problem, coef: np.ndarray, v, dt: float = 0.0,
             aprev: Optional[np.ndarray] = None):
    """CSR matrix and RHS of the steady (dt <= 0, ``:436-1296``) or transient
    (``:1452-2191``) system at frame velocity v."""
    cx, cy, cz = p.cells
    n = ndims(p.cells)
    nz = nnz(p.cells)
    ptr = np.empty(n + 1, dtype=np.int64)
    col = np.empty(nz, dtype=np.int32)
    val = np.empty(nz, dtype=np.float64)
    b = np.empty(n, dtype=np.float64)
    vv = np.ascontiguousarray(v, dtype=np.float64)
    bc = np.ascontiguousarray(p.bc, dtype=np.float64)
    if dt > 0:
        if aprev is None:
            raise ValueError("transient assembly needs the previous A")
        ap = np.ascontiguousarray(aprev[:3], dtype=np.float64)
        app = _p(ap)
    else:
        ap = None
        app = None
    got = lib().qo_assemble(cx, cy, cz, *map(float, p.h), _p(coef), _p(bc), _p(vv), float(dt),
                            app, _p(ptr, _i64), _p(col, _i32), _p(val), _p(b))
    assert got == nz, (got, nz)
    M = sp.csr_matrix((val, col, ptr), shape=(n, n))
    return M, b


def bc_fields(p: Problem, v) -> np.ndarray:
    """Node-sized fields holding the Dirichlet values on the shell, 0 inside."""
    cx, cy, cz = p.cells
    Nx, Ny, Nz = p.nodes
    out = np.zeros((4, Nz, Ny, Nx), dtype=np.float64)
    vv = np.ascontiguousarray(v, dtype=np.float64)
    bc = np.ascontiguousarray(p.bc, dtype=np.float64)
    lib().qo_bc_fields(cx, cy, cz, *map(float, p.h), _p(bc), _p(vv), _p(out))
    return out


def unpack(p: Problem, x: np.ndarray) -> np.ndarray:
    Nx, Ny, Nz = p.nodes
    out = np.empty((4, Nz, Ny, Nx), dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    lib().qo_unpack(*p.cells, _p(xx), _p(out))
    return out


def pack(p: Problem, fields: np.ndarray) -> np.ndarray:
    x = np.empty(ndims(p.cells), dtype=np.float64)
    f = np.ascontiguousarray(fields, dtype=np.float64)
    lib().qo_pack(*p.cells, _p(f), _p(x))
    return x


def hfield(p: Problem, coef: np.ndarray, fields: np.ndarray) -> np.ndarray:
    """(3, cz, cy, cx) cell-centred H (``:1346-1376``)."""
    cx, cy, cz = p.cells
    out = np.empty((3, cz, cy, cx), dtype=np.float64)
    f = np.ascontiguousarray(fields, dtype=np.float64)
    lib().qo_hfield(cx, cy, cz, *map(float, p.h), _p(coef), _p(f), _p(out))
    return out


def bfield(p: Problem, fields: np.ndarray) -> np.ndarray:
    """Cell-centred B: the formula of ``:1353-1372`` with every 1/Mu equal to 1
    (the reference writes only H; B is an additional output of the library)."""
    Nx, Ny, Nz = p.nodes
    ones = np.ones((6, Nz, Ny, Nx), dtype=np.float64)
    return hfield(p, ones, fields)


def trajectory(v0, a, n_a, n, dt) -> float:
    return float(lib().qo_trajectory(v0, a, int(n_a), int(n), dt))


def step_velocity(p: Problem, counter: int):
    """Frame velocity of transient step ``counter`` (``:1431-1449``)."""
    return np.array([lib().qo_step_velocity(float(p.traj[a, 0]), float(p.traj[a, 1]),
                                            int(p.traj[a, 2]), int(counter), float(p.dt))
                     for a in range(3)])


def steady_velocity(p: Problem):
    return np.array([trajectory(p.traj[a, 0], p.traj[a, 1], p.traj[a, 2], 0, p.dt)
                     for a in range(3)])


class DirectSolver:
    """FP64 sparse LU of the row-equilibrated system (stand-in for PARDISO
    phases 13/23/33, ``QMS_3D.f90:1314-1317, 2193-2203``).

    Rows are divided by |diag| first: the raw rows span ~1e-3..1e15, and
    SuperLU's partial pivoting on the raw matrix leaves scaled residuals near
    1e-5.  After equilibration a minimum-degree ordering on A+A^T with
    near-diagonal pivoting reaches ~1e-16.  Two steps of FP64 iterative
    refinement with the same factors follow every solve."""

    def __init__(self, M):
        self.M = M.tocsr()
        self.d = np.abs(self.M.diagonal())
        A = sp.diags(1.0 / self.d) @ self.M
        self.lu = spla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A",
                            diag_pivot_thresh=1e-3,
                            options=dict(SymmetricMode=True))

    def solve(self, b, refine: int = 2):
        bs = b / self.d
        x = self.lu.solve(bs)
        for _ in range(refine):
            r = bs - (self.M @ x) / self.d
            x = x + self.lu.solve(r)
        return x

    def scaled_residual(self, x, b) -> float:
        return float(np.linalg.norm((b - self.M @ x) / self.d)
                     / max(np.linalg.norm(b / self.d), 1e-300))


def direct_solve(M, b):
    return DirectSolver(M).solve(b)


@dataclass
class RunResult:
    fields: List[np.ndarray] = field(default_factory=list)   # (4, Nz, Ny, Nx) per saved state
    H: List[np.ndarray] = field(default_factory=list)        # (3, cz, cy, cx)
    velocity: List[np.ndarray] = field(default_factory=list)
    residual: List[float] = field(default_factory=list)      # ||b - Mx|| / ||b||


def run(p: Problem, steps: Optional[int] = None, keep_every: int = 1) -> RunResult:
    """Steady solve followed by ``steps`` transient solves, exactly the
    program flow of ``QMS_3D.f90:428-2302``.  State 0 is the steady state."""
    coef = coeffs(p)
    nsteps = p.steps if steps is None else steps
    out = RunResult()
    v = steady_velocity(p)
    M, b = assemble(p, coef, v, 0.0)
    S = DirectSolver(M)
    x = S.solve(b)
    out.residual.append(S.scaled_residual(x, b))
    f = unpack(p, x)
    out.fields.append(f)
    out.H.append(hfield(p, coef, f))
    out.velocity.append(v)
    S = None
    vprev = None
    for n in range(1, nsteps + 1):
        v = step_velocity(p, n)
        M, b = assemble(p, coef, v, p.dt, f)
        # refactor only when the velocity changes (phase 23 vs 33, :2193)
        if S is None or vprev is None or not np.array_equal(v, vprev):
            S = DirectSolver(M)
        else:
            S.M = M.tocsr()
        vprev = v
        x = S.solve(b)
        f = unpack(p, x)
        if n % keep_every == 0 or n == nsteps:
            out.fields.append(f)
            out.H.append(hfield(p, coef, f))
            out.velocity.append(v)
            out.residual.append(S.scaled_residual(x, b))
    return out
