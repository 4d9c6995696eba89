"""Batched operator (many designs, one persistent launch) through the C ABI.

Bar: every design's batched output is bit-identical to its single-design
operator launch (the same arithmetic on the same data), whatever the
persistent grid size, the schedule's z-spans and the z-chunk depths, and the
single-design output is
itself checked against the oracle's CSR (1e-13 row-scaled) for one design of
the batch here and for all operator cases in test_gpu_parity."""
import numpy as np
import pytest

from alpineqms_amd import configs, native
from alpineqms_amd.qms_io import material_slots
from oracle import oracle as O
from tests.helpers import problem

pytestmark = pytest.mark.gpu


def _solver(d):
    slots, props = material_slots(d.inputs, d.geom)
    inp = d.inputs
    return native.Solver(inp.cells, inp.h, inp.dt, inp.bc_vector(), slots, props, restart=1)


def _designs():
    # uneven shapes from the batched-design generator, shrunk to test size
    ds = configs.config4_designs(5, steps=2, sizes=[24, 32, 40])
    ds.append(configs.small_box_design(0, 31, nbox=3, v=1.0, ramp=False, gradient=True, cells=(45, 29, 37)))
    return ds


@pytest.fixture(scope="module")
def batch_setup():
    ds = _designs()
    ss = [_solver(d) for d in ds]
    rng = np.random.default_rng(41)
    xs = []
    for s in ss:
        nx, ny, nz = s.nodes
        x = rng.normal(size=(4, nz, ny, nx))
        s.set_fields(x)
        xs.append(x)
    v = rng.normal(size=(len(ss), 3)) * 2.0
    yield ds, ss, xs, v
    for s in ss:
        s.close()


def _single(ss, xs, v, transient):
    out = []
    for s, x, vd in zip(ss, xs, v):
        s.apply(vd, s.dt if transient else 0.0, x)
        out.append(s.get_work())
    return out


@pytest.mark.parametrize("transient", [True, False], ids=["transient", "steady"])
@pytest.mark.parametrize("blocks", [0, 8, 24, 1024, 4096])
def test_batch_bit_identical_to_single(batch_setup, transient, blocks):
    ds, ss, xs, v = batch_setup
    ref = _single(ss, xs, v, transient)
    with native.Batch(ss) as b:
        b.set_blocks(blocks)
        b.apply(v, transient=transient)
        for k, s in enumerate(ss):
            assert np.array_equal(s.get_work(), ref[k]), (k, ds[k].name)


def test_batch_mixed_zchunks_and_tuner(batch_setup):
    ds, ss, xs, v = batch_setup
    for k, s in enumerate(ss):
        s.set_zchunk((1, 2, 3, 5, 7, 64)[k % 6])
        s.set_colwidth((32, 64)[k % 2])
        s.use_coef_classes(k % 3 != 1)          # some designs stream their coefficient arrays
    ref = _single(ss, xs, v, True)
    with native.Batch(ss) as b:
        b.apply(v)
        for k, s in enumerate(ss):
            assert np.array_equal(s.get_work(), ref[k]), k
        kz, cw, blocks = b.tune(v, reps=2)
        assert all(c in (1, 2, 3, 4, 6, 8, 16, 32, 64) for c in kz) and all(c in (32, 64) for c in cw)
        assert blocks >= 0 and blocks % 8 == 0
        b.apply(v)
        for k, s in enumerate(ss):
            assert np.array_equal(s.get_work(), ref[k]), k
        ms_b, by = b.time(v, reps=3)
        ms_s, by2 = b.time(v, reps=3, sequential=True)
        assert ms_b > 0 and ms_s > 0 and by == by2 > 0
    for s in ss:
        s.use_coef_classes(True)


@pytest.mark.parametrize("width,classes", [(32, True), (32, False), (64, True), (64, False)])
def test_batch_uniform_shape_bit_identical(batch_setup, width, classes):
    """Designs sharing one column width and coefficient form run in the
    batch kernel compiled for that variant alone; its rows are the same bits
    as the single-design launches."""
    ds, ss, xs, v = batch_setup
    for k, s in enumerate(ss):
        s.set_colwidth(width)
        s.set_zchunk((3, 16)[k % 2])
        s.use_coef_classes(classes)
    try:
        ref = _single(ss, xs, v, True)
        with native.Batch(ss) as b:
            for blocks in (0, 64, 1024):
                b.set_blocks(blocks)
                b.apply(v)
                for k, s in enumerate(ss):
                    assert np.array_equal(s.get_work(), ref[k]), (width, classes, blocks, k)
    finally:
        for s in ss:
            s.set_colwidth(64)
            s.set_zchunk(4)
            s.use_coef_classes(True)


def test_batch_member_matches_oracle(batch_setup):
    """The last design of the batch (ragged 46 x 30 x 38 nodes) against the
    oracle's CSR: D^-1 M x on free rows, x on fixed rows, 0 where no
    unknown exists."""
    ds, ss, xs, v = batch_setup
    d, s, x, vd = ds[-1], ss[-1], xs[-1], v[-1]
    with native.Batch(ss) as b:
        b.apply(v, transient=True)
    w = s.get_work()
    p = problem(d)
    coef = O.coeffs(p)
    M, _ = O.assemble(p, coef, vd, d.inputs.dt, np.zeros_like(x))
    xp = O.pack(p, x)
    Dg = M.diagonal()
    ref = (M @ xp) / Dg
    got = O.pack(p, w)
    scale = (abs(M) @ np.abs(xp)) / np.abs(Dg)
    assert np.max(np.abs(got - ref) / scale) < 1e-13


def test_batch_of_smallest_and_thin_designs():
    """Designs at the smallest size the reference accepts (5 cells per axis)
    and slabs one column wide or a few planes deep, batched with one larger
    design: z-chunks deeper than a design, segments of one or two planes,
    and most of the persistent workgroups idle.  Every design's rows are
    the same bits as its own launch, for every grid size."""
    shapes = [(5, 5, 5), (5, 31, 6), (40, 5, 5), (6, 7, 70), (45, 29, 37)]
    ds = [configs.small_box_design(0, 7 + k, nbox=1, v=1.0, ramp=False, gradient=True, cells=c)
          for k, c in enumerate(shapes)]
    ss = [_solver(d) for d in ds]
    rng = np.random.default_rng(43)
    try:
        xs = []
        for s in ss:
            nx, ny, nz = s.nodes
            x = rng.normal(size=(4, nz, ny, nx))
            s.set_fields(x)
            xs.append(x)
        v = rng.normal(size=(len(ss), 3))
        for kz in (1, 3, 64):
            for s in ss:
                s.set_colwidth(32)
                s.set_zchunk(kz)
            ref = _single(ss, xs, v, True)
            with native.Batch(ss) as b:
                for blocks in (0, 8, 40, 1024):
                    b.set_blocks(blocks)
                    b.apply(v)
                    for k, s in enumerate(ss):
                        assert np.array_equal(s.get_work(), ref[k]), (kz, blocks, shapes[k])
    finally:
        for s in ss:
            s.close()
