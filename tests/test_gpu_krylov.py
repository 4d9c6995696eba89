"""The Gram-Schmidt step of the solver's FGMRES on the GPU (multi-dot,
two-stage reduction, fused update and norm), through the C ABI, against the
same step in numpy FP64.

Bar: dot products within 1e-13 of |V_i| |w| (their Cauchy-Schwarz scale),
the updated w within 1e-13 of |w| + sum |h_i| |V_i|, its squared norm within
1e-12 relative; and bitwise-reproducible results (the reductions use a fixed
block count and no atomics)."""
import numpy as np
import pytest

from alpineqms_amd import configs, native
from tests.test_gpu_parity import solver_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nv", [1, 5, 8, 9, 31])
def test_gram_schmidt_matches_numpy(nv):
    d = configs.small_box_design(14, 7, nbox=2, ramp=False)
    with solver_for(d) as s:
        nx, ny, nz = s.nodes
        n = 4 * nx * ny * nz
        rng = np.random.default_rng(nv)
        basis = rng.normal(size=(nv, n)) * rng.uniform(0.1, 10.0, size=(nv, 1))
        w = rng.normal(size=n) * 3.0
        h, w2 = s.gram_schmidt(basis, w)
        h_ref = basis @ w
        scale = np.linalg.norm(basis, axis=1) * np.linalg.norm(w)
        assert np.all(np.abs(h[:nv] - h_ref) <= 1e-13 * scale)
        w_ref = w - h[:nv] @ basis          # the update uses the device's dot products
        bound = 1e-13 * (np.abs(w) + np.abs(h[:nv]) @ np.abs(basis))
        assert np.all(np.abs(w2 - w_ref) <= bound)
        assert h[nv] == pytest.approx(float(w2 @ w2), rel=1e-12)
        # fixed reduction order: the same inputs give the same bits
        h2, w3 = s.gram_schmidt(basis, w)
        assert np.array_equal(h, h2) and np.array_equal(w2, w3)


def test_gram_schmidt_orthogonalises_against_an_orthonormal_basis():
    d = configs.small_box_design(12, 2, nbox=2, ramp=False)
    with solver_for(d) as s:
        nx, ny, nz = s.nodes
        n = 4 * nx * ny * nz
        rng = np.random.default_rng(0)
        q, _ = np.linalg.qr(rng.normal(size=(n, 6)))
        w = rng.normal(size=n)
        _, w2 = s.gram_schmidt(np.ascontiguousarray(q.T), w)
        # one classical pass leaves w orthogonal to the basis to round-off
        assert np.max(np.abs(q.T @ w2)) <= 1e-13 * np.linalg.norm(w)


def test_gram_schmidt_rejects_more_vectors_than_the_basis_holds():
    d = configs.small_box_design(12, 2, nbox=2, ramp=False)
    with solver_for(d) as s:
        nx, ny, nz = s.nodes
        n = 4 * nx * ny * nz
        with pytest.raises(native.QmsError):
            s.gram_schmidt(np.zeros((32, n)), np.zeros(n))       # restart 30: at most 31
