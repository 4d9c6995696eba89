"""The batched launch's schedule (host logic, no GPU call): every design's
columns are covered plane by plane exactly once, the persistent workgroups
get about equal work, and the workgroups of one XCD start on neighbouring
columns.  Cost model as in the library: a segment costs its planes plus 2
(the start of its z-march)."""
import math

import numpy as np
import pytest

from alpineqms_amd import configs, native

SEG_START = 2
MIN_SPAN = 2


def _config4_shapes(width=32, band=16):
    ncols, nz = [], []
    for d in configs.config4_designs():
        nx, ny, n3 = d.nodes
        ncols.append(math.ceil(nx / width) * math.ceil(ny / (256 // width)))
        nz.append(n3)
    return ncols, nz, [band] * len(nz)


def _random_shapes(seed, n):
    rng = np.random.default_rng(seed)
    return (list(rng.integers(1, 40, n)), list(rng.integers(6, 130, n)), list(rng.choice([1, 2, 3, 5, 8, 16, 64], n)))


def _natural(ncols, nz, band):
    out = []
    for d, (nc, n3, b) in enumerate(zip(ncols, nz, band)):
        for kb in range(0, n3, b):
            for c in range(nc):
                out.append((d, c, kb, min(kb + b, n3)))
    return out


def _cover(segs, ncols, nz):
    cov = [np.zeros((nc, n3), dtype=np.int32) for nc, n3 in zip(ncols, nz)]
    for d, c, kb, ke in segs:
        assert 0 <= kb < ke <= nz[d] and c < ncols[d]
        cov[d][c, kb:ke] += 1
    return cov


def _loads(segs, off):
    cost = segs[:, 3].astype(np.int64) - segs[:, 2].astype(np.int64) + SEG_START
    return np.array([cost[off[s]:off[s + 1]].sum() for s in range(len(off) - 1)])


@pytest.mark.parametrize("blocks", [0, 8, 24, 1024, 4096])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_every_plane_of_every_column_exactly_once(seed, blocks):
    ncols, nz, band = _random_shapes(seed, 7)
    segs, off = native.batch_plan(ncols, nz, band, blocks)
    nslots = len(off) - 1
    assert nslots % 8 == 0 and nslots >= 8
    assert off[0] == 0 and off[-1] == len(segs) and np.all(np.diff(off.astype(np.int64)) >= 0)
    for c in _cover(segs, ncols, nz):
        assert np.all(c == 1)
    if blocks:
        assert nslots == (blocks + 7) // 8 * 8


def test_one_workgroup_per_natural_piece():
    ncols, nz, band = _random_shapes(5, 5)
    segs, off = native.batch_plan(ncols, nz, band, 0)
    nat = _natural(ncols, nz, band)
    assert [tuple(int(v) for v in s) for s in segs] == nat
    assert len(off) - 1 == max(8, (len(nat) + 7) // 8 * 8)
    assert np.all(np.diff(off.astype(np.int64)) <= 1)


@pytest.mark.parametrize("band", [8, 16, 32, 64])
def test_persistent_slots_balanced_on_config4(band):
    ncols, nz, bands = _config4_shapes(band=band)
    segs, off = native.batch_plan(ncols, nz, bands, 1024)
    per = 1024 // 8
    loads = _loads(segs, off)
    for x in range(8):
        lx = loads[x * per:(x + 1) * per]
        # each slot within a split's worth of its XCD's mean
        assert lx.max() <= math.ceil(lx.mean()) + 2 * MIN_SPAN + SEG_START
    assert loads.max() / loads.mean() < 1.08
    # splitting adds at most about one segment per slot
    assert len(segs) <= len(_natural(ncols, nz, bands)) + 2 * 1024


def test_xcd_slots_start_on_consecutive_pieces():
    """The first segments of an XCD's slots are consecutive natural pieces:
    the workgroups that run together on that XCD work on neighbouring
    columns of one design at the same depth."""
    ncols, nz, band = _config4_shapes(band=16)
    segs, off = native.batch_plan(ncols, nz, band, 1024)
    nat = _natural(ncols, nz, band)
    index = {p[:3]: i for i, p in enumerate(nat)}
    per = 128
    for x in range(8):
        firsts = [index[tuple(int(v) for v in segs[off[x * per + l]][:3])] for l in range(per)]
        assert firsts == list(range(firsts[0], firsts[0] + per))


def test_plan_rejects_bad_input():
    with pytest.raises(native.QmsError):
        native.batch_plan([4, 0], [10, 10], [4, 4], 64)
    with pytest.raises(native.QmsError):
        native.batch_plan([4], [10], [4], -1)


def test_plan_properties_hold_for_any_batch():
    """Property check over generated batches: exact cover of every column's
    planes, grid size a multiple of 8, and, with persistent workgroups, no
    workgroup more than a split's cost above its XCD's mean."""
    hyp = pytest.importorskip("hypothesis")
    st = hyp.strategies

    design = st.tuples(st.integers(1, 30), st.integers(6, 90), st.sampled_from([1, 2, 3, 4, 7, 8, 16, 32, 64]))

    @hyp.settings(max_examples=60, deadline=None)
    @hyp.given(st.lists(design, min_size=1, max_size=9), st.sampled_from([0, 8, 16, 40, 256, 1024]))
    def check(ds, blocks):
        ncols, nz, band = (list(v) for v in zip(*ds))
        segs, off = native.batch_plan(ncols, nz, band, blocks)
        nslots = len(off) - 1
        assert nslots % 8 == 0 and off[-1] == len(segs)
        for c in _cover(segs, ncols, nz):
            assert np.all(c == 1)
        if blocks:
            per = nslots // 8
            loads = _loads(segs, off)
            for x in range(8):
                lx = loads[x * per:(x + 1) * per]
                if lx.sum():
                    assert lx.max() <= math.ceil(lx.mean()) + 2 * MIN_SPAN + SEG_START

    check()
