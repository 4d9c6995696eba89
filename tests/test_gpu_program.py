"""The drop-in executable on the GPU: `QMS_3D inputs.txt` in a run directory
holding the files master.py writes (QMS_3D.f90:120-165, :229-232) writes the
Hx/Hy/Hz.bin slice records and H*_all.bin volumes that plot.py and the
ParaView scripts read (:1379-1404, :2275-2300), with values that match the
oracle's run of the same design to the solve bar."""
import os
import subprocess
import sys

import numpy as np
import pytest

from alpineqms_amd import configs, program
from alpineqms_amd.qms_io import read_records, write_geom, write_inputs
from oracle import oracle as O
from tests.helpers import problem, rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_drop_in_program_writes_the_reference_records(tmp_path, precision):
    d = configs.small_box_design(9, 4, nbox=2, v=3.0, ramp=True, gradient=True, steps=3)
    inp = d.inputs
    inp.slice_axis, inp.slice_loc, inp.tt, inp.save_all = "y", 3, 2, "yes"
    write_inputs(str(tmp_path / "inputs.txt"), inp)
    write_geom(str(tmp_path / inp.geom_name), d.geom)
    env = dict(os.environ, QMS_PRECISION=precision)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "alpineqms_amd", "QMS_3D"), "inputs.txt"],
                       cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "Main time marching completed" in r.stdout
    ref = O.run(problem(d))
    vols = {n: read_records(str(tmp_path / f"{n}_all.bin")) for n in ("Hx", "Hy", "Hz")}
    slices = {n: read_records(str(tmp_path / f"{n}.bin")) for n in ("Hx", "Hy", "Hz")}
    assert all(len(v) == 1 + inp.steps for v in vols.values())
    assert all(len(v) == 1 + inp.steps // inp.tt for v in slices.values())
    cx, cy, cz = inp.cells
    for n, Hr in enumerate(ref.H):
        H = np.stack([vols[c][n].reshape(cz, cy, cx) for c in ("Hx", "Hy", "Hz")])
        assert rel(H, Hr) < 1e-8, (n, rel(H, Hr))
    for s, n in enumerate(range(0, inp.steps + 1, inp.tt)):
        S = program._slice(ref.H[n], "y", 3, inp.cells)
        got = np.stack([slices[c][s].reshape(S.shape[1:]) for c in ("Hx", "Hy", "Hz")])
        assert rel(got, S) < 1e-8
