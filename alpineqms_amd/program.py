"""The QMS_3D program flow on the GPU: a drop-in for ``./QMS_3D inputs.txt``.

Mirrors the main program of ``3D/QMS_3D.f90``: read ``inputs.txt`` and the
geometry (``:117-165, :229-254``), steady-state initialisation
(``:424-1415``), then ``time_steps`` implicit steps (``:1424-2302``), writing
the same ``Hx/Hy/Hz.bin`` slice records and ``H*_all.bin`` volume records.
The assemble-and-PARDISO part of every solve runs in libqms.so.

Options that the reference grammar has no place for come from the
environment, so ``inputs.txt`` is never extended:

* ``QMS_PRECISION``: ``fp64`` (default) or ``mixed`` (FP32 multigrid inside
  FP64 Krylov);
* ``QMS_TOL``: relative scaled residual of each solve (default 1e-12);
* ``QMS_EXTRA_OUTPUTS=1``: also write ``Ax/Ay/Az_all.bin``, ``V_all.bin`` and
  ``Bx/By/Bz_all.bin`` in the same record format.
"""
from __future__ import annotations

import os
import struct
import sys
import time
from typing import Optional

import numpy as np

from . import native
from .qms_io import InputsError, RecordWriter, material_slots, read_geom, read_inputs

# Default-real literal 1E-8 of the first-step nudge (QMS_3D.f90:1432-1439).
NUDGE = float(np.float32(1e-8))


def trajectory_1(v0: float, a: float, n_a: int, n: int, dt: float) -> float:
    """``trajectory_1``, QMS_3D.f90:2331-2350."""
    if n <= n_a:
        return v0 + a * n * dt
    return v0 + a * n_a * dt


def steady_velocity(traj, dt: float) -> np.ndarray:
    return np.array([trajectory_1(t[0], t[1], t[2], 0, dt) for t in traj])


def step_velocity(traj, dt: float, counter: int) -> np.ndarray:
    """Velocity of transient step ``counter`` including the first-step nudge
    (QMS_3D.f90:1431-1449)."""
    out = []
    for (v0, a, n_a) in traj:
        nudge = 0.0
        if counter == 1:
            vs = trajectory_1(v0, a, n_a, 0, dt)
            if a * dt < v0 * NUDGE:
                nudge = vs * NUDGE / dt
        out.append(trajectory_1(v0, a + nudge, n_a, counter, dt))
    return np.array(out)


def precision_from_env() -> int:
    p = os.environ.get("QMS_PRECISION", "fp64").lower()
    if p in ("fp64", "double", "0"):
        return native.QMS_FP64
    if p in ("mixed", "fp32", "1"):
        return native.QMS_MIXED
    raise InputsError(f"QMS_PRECISION must be fp64 or mixed, not {p!r}")


def make_solver(inp, geom, precision: Optional[int] = None, **kw) -> native.Solver:
    slots, props = material_slots(inp, geom)
    if precision is None:
        precision = precision_from_env()
    tol = float(os.environ.get("QMS_TOL", kw.pop("tol", 1e-12)))
    return native.Solver(inp.cells, inp.h, inp.dt, inp.bc_vector(), slots, props,
                         precision=precision, tol=tol, **kw)


def _slice(H: np.ndarray, axis: str, loc: int, cells) -> np.ndarray:
    """Hx_out etc., QMS_3D.f90:1380-1394 (H is (3, cz, cy, cx))."""
    cx, cy, cz = cells
    lim = {"x": cx, "y": cy, "z": cz}[axis]
    if not 1 <= loc <= lim:
        raise InputsError(f"slice location {loc} outside 1..{lim}")
    if axis == "x":
        return H[:, :, :, loc - 1]
    if axis == "y":
        return H[:, :, loc - 1, :]
    return H[:, loc - 1, :, :]


def _fmt_seconds(t: float) -> str:
    s = f"{t:6.2f}"
    return s if len(s) <= 6 else "******"       # Fortran F6.2 overflow


def main(argv=None) -> int:
    argv = sys.argv if argv is None else argv
    t0 = time.time()
    print("Reading in the inputs file and setting up the simulation...", flush=True)
    if len(argv) < 2:
        print("usage: QMS_3D <inputs-file>", file=sys.stderr)
        return 2
    inp = read_inputs(argv[1][:50])
    if inp.slice_axis not in ("x", "y", "z"):
        raise InputsError(f"output slice axis must be x, y or z, not {inp.slice_axis!r}")
    if inp.tt < 1:
        raise InputsError("save interval must be at least 1")
    geom = read_geom(inp.geom_name, inp.cells)
    solver = make_solver(inp, geom)
    extra = os.environ.get("QMS_EXTRA_OUTPUTS", "0") == "1"
    save_all = inp.save_all.strip() == "yes"
    outs = {n: RecordWriter(f"{n}.bin") for n in ("Hx", "Hy", "Hz")}
    if save_all:
        outs.update({f"{n}_all": RecordWriter(f"{n}_all.bin") for n in ("Hx", "Hy", "Hz")})
    if extra:
        outs.update({f"{n}_all": RecordWriter(f"{n}_all.bin") for n in ("Ax", "Ay", "Az", "V", "Bx", "By", "Bz")})
    cx, cy, cz = inp.cells

    def emit(counter: int):
        H = solver.get_h()
        if counter % inp.tt == 0:
            S = _slice(H, inp.slice_axis, inp.slice_loc, inp.cells)
            for c, n in enumerate(("Hx", "Hy", "Hz")):
                outs[n].write(S[c])
        if save_all:
            for c, n in enumerate(("Hx", "Hy", "Hz")):
                outs[f"{n}_all"].write(H[c])
        if extra:
            f = solver.get_fields()
            outs["Ax_all"].write(f[0][:, :, :cx])
            outs["Ay_all"].write(f[1][:, :cy, :])
            outs["Az_all"].write(f[2][:cz, :, :])
            outs["V_all"].write(f[3])
            B = solver.get_h(unit_mu=True)
            for c, n in enumerate(("Bx", "By", "Bz")):
                outs[f"{n}_all"].write(B[c])

    t1 = time.time()
    print(f"Setup completed in {_fmt_seconds(t1 - t0)} seconds.", flush=True)
    print("Intializing steady state...", flush=True)
    solver.steady(steady_velocity(inp.traj, inp.dt))
    emit(0)
    t2 = time.time()
    print(f"Initialization of steady state completed in {_fmt_seconds(t2 - t1)} seconds.", flush=True)
    print("Starting main time marching algorithm...", flush=True)
    for counter in range(1, inp.steps + 1):
        solver.step(step_velocity(inp.traj, inp.dt, counter))
        if counter % 10 == 0:
            print(f"{counter} of {inp.steps} time steps", flush=True)
        emit(counter)
    for w in outs.values():
        w.close()
    solver.close()
    print(f"Main time marching completed in {_fmt_seconds(time.time() - t2)} seconds.", flush=True)
    return 0


def cli() -> int:
    try:
        return main()
    except (InputsError, native.QmsError) as e:
        print(str(e), file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(cli())
