"""Many designs on one GPU (configs 4 and 5): the asynchronous run, the
result of a design alone against the same design served among others, the
device-footprint estimate the admission uses, and admission under a memory
budget."""
import numpy as np
import pytest

from alpineqms_amd import configs, native, program, serve
from alpineqms_amd.configs import Request

pytestmark = pytest.mark.gpu


def _design(seed, n=14, steps=3, v=3.0):
    return configs.small_box_design(n, seed, nbox=2, v=v, ramp=True, gradient=True, steps=steps)


def _run_alone(d, precision):
    with serve.make_solver(d, precision) as s:
        s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
        s.run(serve.velocities(d, 1, d.inputs.steps))
        return s.get_fields()


@pytest.mark.parametrize("precision", [native.QMS_FP64, native.QMS_MIXED], ids=["fp64", "mixed"])
def test_async_run_matches_blocking_run(precision):
    d = _design(31)
    ref = _run_alone(d, precision)
    with serve.make_solver(d, precision) as s:
        s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
        s.run_async(serve.velocities(d, 1, d.inputs.steps))
        with pytest.raises(native.QmsError):
            s.get_fields()                  # the handle is busy while its run goes on
        s.wait()
        assert not s.poll()
        assert np.array_equal(s.get_fields(), ref)


@pytest.mark.parametrize("precision", [native.QMS_FP64, native.QMS_MIXED], ids=["fp64", "mixed"])
def test_design_alone_and_in_the_mix_give_the_same_result(precision):
    """Served among five other designs (all in flight together, their kernels
    overlapping on the GPU), a design's final state is bit for bit its state
    when it runs alone: every reduction is deterministic per design."""
    ds = [_design(40 + k, n=12 + 2 * k, steps=3 + k % 3, v=2.0 + k) for k in range(6)]
    alone = {d.name: _run_alone(d, precision) for d in ds}
    got = {}

    def keep(req, s):
        got[req.design.name] = s.get_fields()

    res = serve.serve([Request(0.0, d) for d in ds], precision=precision, max_active=6, on_result=keep)
    assert all(r.error is None for r in res.records), [r.error for r in res.records]
    assert res.peak_active > 1
    for d in ds:
        assert np.array_equal(got[d.name], alone[d.name]), d.name


def test_footprint_estimate_covers_the_handle():
    for prec in (native.QMS_FP64, native.QMS_MIXED):
        d = _design(50, n=24)
        est = serve.footprint(d, prec)
        with serve.make_solver(d, prec) as s:
            s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
            s.step(program.step_velocity(d.inputs.traj, d.inputs.dt, 1))
            held = s.stats().device_bytes
        assert held <= est <= 1.6 * held, (prec, held, est)


def test_admission_keeps_requests_in_flight_under_the_budget():
    ds = [_design(60 + k, n=16, steps=2) for k in range(5)]
    one = serve.footprint(ds[0], native.QMS_MIXED)
    budget = int(2.5 * one)                 # room for two requests at a time
    reqs = [Request(0.01 * k, d) for k, d in enumerate(ds)]
    res = serve.serve(reqs, precision=native.QMS_MIXED, budget_bytes=budget, max_active=8)
    assert all(r.error is None for r in res.records)
    assert res.peak_bytes <= budget and res.peak_active <= 2
    assert all(r.end >= r.start >= r.arrival for r in res.records)
    # a request larger than the whole budget is refused, not run
    big = Request(0.0, _design(70, n=40))
    res = serve.serve([big], precision=native.QMS_MIXED, budget_bytes=budget)
    assert res.records[0].error and "exceeds" in res.records[0].error
