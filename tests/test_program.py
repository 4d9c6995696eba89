"""The drop-in program end to end on the CPU: `QMS_3D inputs.txt` reads the
files master.py writes and writes the records plot.py reads.

The GPU handle is replaced by a stand-in that solves each step with the
oracle's direct solver, so this pins the host side of the boundary -- the
velocity passed to every solve (QMS_3D.f90:428-433, :1431-1449), the record
sequence and shapes of Hx/Hy/Hz.bin and H*_all.bin (:1379-1404, :2275-2300;
counts 1 + floor(Nt/tt) and 1 + Nt) and the slice indexing (:1380-1394) --
against the oracle's own run of the same design."""
import numpy as np
import pytest

from alpineqms_amd import configs, program
from alpineqms_amd.qms_io import read_records, write_geom, write_inputs
from oracle import oracle as O
from tests.helpers import problem


class OracleHandle:
    """Stand-in for native.Solver: the same calls, answered by the oracle."""

    def __init__(self, p):
        self.p = p
        self.coef = O.coeffs(p)
        self.f = None
        self.velocities = []

    def steady(self, v):
        self.velocities.append(np.array(v, dtype=float))
        M, b = O.assemble(self.p, self.coef, v, 0.0)
        self.f = O.unpack(self.p, O.DirectSolver(M).solve(b))

    def step(self, v):
        self.velocities.append(np.array(v, dtype=float))
        M, b = O.assemble(self.p, self.coef, v, self.p.dt, self.f)
        self.f = O.unpack(self.p, O.DirectSolver(M).solve(b))

    def get_h(self, unit_mu=False):
        return O.bfield(self.p, self.f) if unit_mu else O.hfield(self.p, self.coef, self.f)

    def get_fields(self):
        return self.f.copy()

    def close(self):
        pass


@pytest.mark.parametrize("axis,loc,tt,steps", [("y", 3, 2, 5), ("x", 7, 1, 3), ("z", 1, 4, 4)])
def test_program_records_match_oracle_run(tmp_path, monkeypatch, axis, loc, tt, steps):
    d = configs.small_box_design(9, 4, nbox=2, v=3.0, ramp=True, gradient=True, steps=steps)
    inp = d.inputs
    inp.slice_axis, inp.slice_loc, inp.tt, inp.save_all = axis, loc, tt, "yes"
    monkeypatch.chdir(tmp_path)
    write_inputs("inputs.txt", inp)
    write_geom(inp.geom_name, d.geom)
    handles = []

    def make(inp_, geom_, precision=None, **kw):
        # the program re-reads both files: what it passes on is what it parsed
        assert inp_ == inp and np.array_equal(geom_, d.geom)
        h = OracleHandle(problem(d))
        handles.append(h)
        return h

    monkeypatch.setattr(program, "make_solver", make)
    monkeypatch.setenv("QMS_EXTRA_OUTPUTS", "1")
    assert program.main(["QMS_3D", "inputs.txt"]) == 0
    (h,) = handles

    p = problem(d)
    ref = O.run(p, steps=steps)
    # one solve per state, at the oracle's velocities
    assert len(h.velocities) == steps + 1
    assert np.array_equal(h.velocities[0], O.steady_velocity(p))
    for n in range(1, steps + 1):
        assert np.array_equal(h.velocities[n], O.step_velocity(p, n))

    cx, cy, cz = inp.cells
    ai = {"x": 3, "y": 2, "z": 1}[axis]                   # axis of the (3, cz, cy, cx) H array
    saved = [n for n in range(steps + 1) if n % tt == 0]
    for c, name in enumerate(("Hx", "Hy", "Hz")):
        recs = read_records(f"{name}.bin")
        assert len(recs) == 1 + steps // tt
        for r, n in zip(recs, saved):
            S = np.take(ref.H[n][c], loc - 1, axis=ai - 1)  # Fortran order of the slice = C order here
            assert r.size == S.size
            np.testing.assert_allclose(r, S.ravel(), rtol=1e-9, atol=1e-9 * np.abs(ref.H[n][c]).max())
        allr = read_records(f"{name}_all.bin")
        assert len(allr) == 1 + steps
        for r, n in zip(allr, range(steps + 1)):
            assert r.size == cx * cy * cz
            np.testing.assert_allclose(r, ref.H[n][c].ravel(), rtol=1e-9, atol=1e-9 * np.abs(ref.H[n][c]).max())
    # build additions: valid sub-box shapes of A, node-sized V, cell B
    assert read_records("Ax_all.bin")[0].size == cx * (cy + 1) * (cz + 1)
    assert read_records("Ay_all.bin")[0].size == (cx + 1) * cy * (cz + 1)
    assert read_records("Az_all.bin")[0].size == (cx + 1) * (cy + 1) * cz
    assert read_records("V_all.bin")[0].size == (cx + 1) * (cy + 1) * (cz + 1)
    assert len(read_records("Bx_all.bin")) == 1 + steps


def test_program_stops_on_bad_slice(tmp_path, monkeypatch):
    d = configs.small_box_design(9, 4, nbox=1, steps=1)
    inp = d.inputs
    inp.slice_axis = "w"
    monkeypatch.chdir(tmp_path)
    write_inputs("inputs.txt", inp)
    write_geom(inp.geom_name, d.geom)
    monkeypatch.setattr(program, "make_solver", lambda *a, **k: pytest.fail("no solver before the check"))
    monkeypatch.setattr(program.sys, "argv", ["QMS_3D", "inputs.txt"])
    assert program.cli() == 1
