"""Host-side logic that needs no GPU: the C-ABI library exports, the QMS_3D
program helpers (velocity schedule, output slices) and the design generators."""
import ctypes
import os
import re

import numpy as np
import pytest

from alpineqms_amd import configs, native, program
from alpineqms_amd.qms_io import InputsError
from oracle import oracle as O
from tests.helpers import problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "qms.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int64_t|int|void|const char \*)\s*(qms_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libqms.so not built")
    lib = ctypes.CDLL(native.LIB_PATH)
    names = header_functions()
    assert names and set(names) == set(native.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_config_defaults_without_gpu():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libqms.so not built")
    c = native.default_config()
    assert (c.precision, c.restart, c.mg_nu) == (native.QMS_FP64, 30, 4)
    assert c.tol == 1e-12 and c.mg_omega == 0.7 and c.mg_alpha == 1.5
    assert (c.mg_nu_coarse, c.mg_nu_v, c.mg_omega_v, c.mg_alpha_v, c.mg_fused_tail) == (3, 6, 0.85, 1.9, 1)
    assert (c.mg_cheb_lmax, c.mg_cheb_ratio, c.mg_cheb_lmax_v, c.mg_cheb_ratio_v) == (2.6, 4.0, 2.0, 8.0)
    assert c.mg_fused_restrict == 1


def test_python_config_defaults_match_the_library():
    """make_config (which every Solver goes through) starts from the C
    defaults and sets every field: its own defaults must be the library's."""
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libqms.so not built")
    c = native.default_config()
    m = native.make_config((8, 8, 8), (1.0, 1.0, 1.0), 1.0, np.zeros(12))
    for f, _ in native.Config._fields_:
        if f in ("cells", "h", "dt", "bc"):
            continue
        assert getattr(c, f) == getattr(m, f), f


@pytest.mark.parametrize("cells,h,restart", [((4, 5, 5), 5e-4, 30), ((5, 5, 5), 0.0, 30), ((5, 5, 5), 5e-4, 0),
                                                ((5, 5, 4097), 5e-4, 30)])
def test_create_rejects_invalid_configs_before_touching_the_gpu(cells, h, restart):
    """qms_create checks its arguments (cells per axis in [5, 4096], positive
    steps, restart in [1, 64]) and returns QMS_ERR_ARG with a message before
    any device call, so this runs without a GPU."""
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libqms.so not built")
    import numpy as np
    nc = cells[0] * cells[1] * cells[2]
    with pytest.raises(native.QmsError) as e:
        native.Solver(cells, (h, h, h), 1e-6, np.zeros(12), np.zeros(nc, dtype=np.uint16),
                      np.array([[1.0, 1.0, 1.0, 0.0, 0.0, 0.0]]), restart=restart)
    assert e.value.code == native.QMS_ERR_ARG and "qms_create" in str(e.value)


@pytest.mark.parametrize("traj", [
    [(5.0, 0.0, 0), (0.0, 0.0, 0), (-1.0, 0.0, 0)],
    [(0.0, 1e5, 40), (2.0, 0.0, 3), (1.0, 3.0, 2)],
    [(2.0, 1e-9, 5), (0.5, 0.0, 1), (0.0, 0.0, 0)],
])
def test_velocity_schedule_matches_oracle(traj):
    """program.step_velocity (the product) against the oracle's C restatement."""
    dt = 2e-5
    p = O.Problem(cells=(5, 5, 5), h=(1, 1, 1), dt=dt, steps=10, bc=np.zeros(12), traj=np.array(traj, float),
                  materials=np.zeros((0, 7)), geom=np.zeros((5, 5, 5)))
    assert np.array_equal(program.steady_velocity(traj, dt), O.steady_velocity(p))
    for n in range(1, 12):
        assert np.array_equal(program.step_velocity(traj, dt, n), O.step_velocity(p, n)), n


def test_slices_match_reference_indexing():
    """Hx_out = Hx(loc, 1:y, 1:z) etc. (QMS_3D.f90:1380-1394), written in
    Fortran order: y fastest for an x slice."""
    cells = (4, 5, 6)
    H = np.arange(3 * 6 * 5 * 4, dtype=float).reshape(3, 6, 5, 4)    # (c, z, y, x)
    Sx = program._slice(H, "x", 2, cells)
    assert Sx.shape == (3, 6, 5) and np.array_equal(Sx[0], H[0][:, :, 1])
    Sy = program._slice(H, "y", 5, cells)
    assert np.array_equal(Sy[1], H[1][:, 4, :])
    Sz = program._slice(H, "z", 1, cells)
    assert np.array_equal(Sz[2], H[2][0])
    with pytest.raises(InputsError):
        program._slice(H, "z", 7, cells)


def test_fortran_f62_overflow():
    assert program._fmt_seconds(3.14159) == "  3.14"
    assert program._fmt_seconds(1234.5) == "******"


def test_precision_from_env(monkeypatch):
    monkeypatch.setenv("QMS_PRECISION", "mixed")
    assert program.precision_from_env() == native.QMS_MIXED
    monkeypatch.setenv("QMS_PRECISION", "fp64")
    assert program.precision_from_env() == native.QMS_FP64
    monkeypatch.setenv("QMS_PRECISION", "half")
    with pytest.raises(InputsError):
        program.precision_from_env()


def test_design_generators_are_seeded_and_in_bounds():
    a, b = configs.config2(N=24, steps=2), configs.config2(N=24, steps=2)
    assert np.array_equal(a.geom, b.geom) and a.inputs == b.inputs
    for d in [configs.config1(), configs.config2(N=32), configs.config3(N=40)] + configs.config4_designs(4):
        p = problem(d)
        ids = {int(m.id) for m in d.inputs.materials}
        assert set(np.unique(d.geom).astype(int)) <= ids
        assert O.coeffs(p).shape[0] == 6
        # boxes keep the 2-layer Dirichlet shell and the index-1 planes clear
        g = d.geom
        assert np.all(g[:2] == 1) and np.all(g[:, :2] == 1) and np.all(g[:, :, :2] == 1)
    c1 = configs.config1()
    assert c1.inputs.cells == (31, 31, 31) and c1.inputs.traj[0] == (-1.0, 0.0, 0)
    assert np.count_nonzero(c1.geom == 2) == 12 ** 3
    r = configs.config3(N=128, ramp=True)
    assert r.inputs.traj[0][2] == 200 and r.inputs.traj[0][1] * 200 * 2e-6 == pytest.approx(5.0)
    s = configs.config5_stream(20, rate=1.0)
    assert all(32 <= q.design.nodes[0] <= 160 and q.design.nodes[0] % 8 == 0 for q in s)
    assert all(s[i].arrival < s[i + 1].arrival for i in range(len(s) - 1))


def test_footprint_estimate_is_host_only_and_grows_with_the_design():
    """qms_estimate_bytes needs no GPU; it grows with the grid and with the
    FP64 multigrid, and refuses a grid the library refuses."""
    from alpineqms_amd import native
    small = native.estimate_bytes((31, 31, 31), native.QMS_FP64)
    big = native.estimate_bytes((127, 127, 127), native.QMS_FP64)
    mixed = native.estimate_bytes((127, 127, 127), native.QMS_MIXED)
    assert 0 < small < big and mixed < big
    # the Krylov basis alone (31 FP64 vectors of 4 N^3) is a lower bound
    assert big > 31 * 8 * 4 * 128 ** 3
    assert native.estimate_bytes((3, 31, 31), native.QMS_FP64) == 0
