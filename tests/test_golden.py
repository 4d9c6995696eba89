"""The committed golden fixtures (tests/golden/make_golden.py) against the
oracle and the design generators: the assembled operator must match bit for
bit, a short run to solver round-off."""
import os

import numpy as np

from oracle import oracle as O
from alpineqms_amd import configs
from tests.helpers import problem

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_csr_fixture_bitwise():
    g = np.load(os.path.join(GOLD, "csr_N8.npz"))
    d = configs.small_box_design(8, 31, nbox=1, v=1.0, ramp=False, gradient=True)
    p = problem(d)
    coef = O.coeffs(p)
    for tag, dt in (("steady", 0.0), ("transient", p.dt)):
        M, b = O.assemble(p, coef, np.array([1.0, -0.5, 0.25]), dt, g["aprev"])
        assert np.array_equal(M.indptr, g[f"{tag}_indptr"])
        assert np.array_equal(M.indices, g[f"{tag}_indices"])
        assert np.array_equal(M.data, g[f"{tag}_data"])
        assert np.array_equal(b, g[f"{tag}_b"])


def test_run_fixture():
    g = np.load(os.path.join(GOLD, "run_config1_ramp_N16.npz"))
    d = configs.config1(16, steps=2, ramp=True)
    r = O.run(problem(d), steps=2)
    assert np.array_equal(np.stack(r.velocity), g["velocity"])
    f, H = np.stack(r.fields), g["H"]
    for n in range(f.shape[0]):
        a, ga = f[n, :3], g["fields"][n, :3]
        assert np.linalg.norm(a - ga) <= 1e-10 * np.linalg.norm(ga)
        assert np.linalg.norm(np.stack(r.H)[n] - H[n]) <= 1e-10 * np.linalg.norm(H[n])
    assert np.all(g["residual"] < 1e-13)
