"""The library's own HIP sources built for the CPU (tests/cpuemu) and run under
AddressSanitizer + UBSan: every kernel launch of the operator, the RHS, H,
the multigrid preconditioner and the FGMRES solve executes with the GPU's
grid and indexing on host memory, so an out-of-bounds access, a missing
barrier or a wrong index fails here before any GPU run.

The bodies run in a child process with the sanitizer runtime preloaded; the
parity bars are the GPU tests' bars against the oracle."""
import os
import subprocess
import sys
import textwrap

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "cpuemu"))


@pytest.fixture(scope="module")
def cpu_lib():
    import build as emu   # tests/cpuemu/build.py
    try:
        path = emu.build()
        asan = emu.asan_runtime()
    except Exception as e:   # no g++ with sanitizers in this environment
        pytest.skip(f"CPU build of the HIP sources unavailable: {e}")
    return path, asan


def run_child(cpu_lib, body: str, timeout: int = 600) -> str:
    path, asan = cpu_lib
    prog = textwrap.dedent(f"""
        import sys
        sys.path.insert(0, {ROOT!r})
        import numpy as np
        from alpineqms_amd import native
        native._lib = native.load({path!r})
    """) + textwrap.dedent(body)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-c", prog], env=env, capture_output=True, text=True, timeout=timeout)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    return out


def test_operator_rhs_h_match_oracle(cpu_lib):
    """The smoke body (fused operator, step RHS, cell-centred H) on the CPU
    build, and the operator on a ragged grid with partial columns."""
    out = run_child(cpu_lib, """
        import __graft_entry__ as g
        g.smoke()
        from tests.test_gpu_parity import _operator_check
        from alpineqms_amd import configs
        d = configs.small_box_design(0, 11, nbox=2, v=2.0, ramp=False, gradient=True, cells=(13, 21, 9))
        _operator_check(d, np.array([-2.0, 0.5, 1.25]), True, (3, 64), seed=13)
        print("ok")
    """)
    assert "ok" in out


@pytest.mark.parametrize("precision", ["QMS_FP64", "QMS_MIXED"])
def test_solve_matches_direct(cpu_lib, precision):
    """Steady state and two ramp steps of a 10-node design with gradients and a
    frame velocity: the solve path (multigrid preconditioner, device-side
    FGMRES) against the oracle's direct solve, to the GPU tests' 1e-8 bar."""
    out = run_child(cpu_lib, f"""
        from alpineqms_amd import configs
        from tests import test_gpu_parity as T
        d = configs.small_box_design(10, 3, nbox=2, v=3.0, ramp=True, gradient=True, steps=2)
        T.check_solve(d, native.{precision}, 2)
        print("ok")
    """)
    assert "ok" in out


def test_fused_multigrid_tail_matches_launches(cpu_lib):
    """The V-cycle's last two levels in one single-workgroup launch
    (mg_fused_tail, 1024 fibers with the tail's LDS) against one launch per
    sweep, both precisions: bit-identical fields after the steady state and
    two ramp steps."""
    out = run_child(cpu_lib, """
        from alpineqms_amd import configs, program
        from tests.test_gpu_parity import solver_for
        d = configs.small_box_design(10, 3, nbox=2, v=3.0, ramp=True, gradient=True, steps=2)
        for prec in (native.QMS_FP64, native.QMS_MIXED):
            runs = []
            for tail in (True, False):
                states = []
                with solver_for(d, precision=prec, mg_fused_tail=tail) as s:
                    s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
                    states.append(s.get_fields())
                    for n in (1, 2):
                        s.step(program.step_velocity(d.inputs.traj, d.inputs.dt, n))
                        states.append(s.get_fields())
                runs.append(states)
            for n, (a, b) in enumerate(zip(*runs)):
                assert np.array_equal(a, b), (prec, n, float(np.max(np.abs(a - b))))
        print("ok")
    """)
    assert "ok" in out


def test_fused_level0_restriction_matches_launches(cpu_lib):
    """Level 0's residual, restriction and level 1's first sweep in one
    z-march launch (mg_fused_restrict) against two launches, both precisions,
    on a ragged design with four levels (level 1 is not the tail's):
    bit-identical fields after the steady state and two ramp steps."""
    out = run_child(cpu_lib, """
        from alpineqms_amd import configs, program
        from tests.test_gpu_parity import solver_for
        d = configs.small_box_design(0, 3, nbox=2, v=3.0, ramp=True, gradient=True, steps=2, cells=(17, 14, 16))
        for prec in (native.QMS_FP64, native.QMS_MIXED):
            runs = []
            for fused in (True, False):
                states = []
                with solver_for(d, precision=prec, mg_fused_restrict=fused) as s:
                    s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
                    states.append(s.get_fields())
                    for n in (1, 2):
                        s.step(program.step_velocity(d.inputs.traj, d.inputs.dt, n))
                        states.append(s.get_fields())
                    assert s.stats().n_levels_a >= 4 and s.stats().n_levels_v >= 4, s.stats().n_levels_a
                    assert s.stats().fused_restrict_blocks == (2 if fused else 0)
                runs.append(states)
            for n, (a, b) in enumerate(zip(*runs)):
                assert np.array_equal(a, b), (prec, n, float(np.max(np.abs(a - b))))
        print("ok")
    """)
    assert "ok" in out


def test_async_runs_and_served_mix_match_runs_alone(cpu_lib):
    """The asynchronous run and the serving loop with three designs in flight
    on the CPU build (library threads, one stream per handle): each design's
    result is bit for bit its result alone, and a handle refuses calls while
    its run goes on."""
    out = run_child(cpu_lib, """
        from alpineqms_amd import serve
        from alpineqms_amd.configs import Request
        import tests.test_gpu_serve as T
        ds = [T._design(80 + k, n=9 + k, steps=2, v=2.0 + k) for k in range(3)]
        alone = {d.name: T._run_alone(d, native.QMS_MIXED) for d in ds}
        got = {}
        res = serve.serve([Request(0.0, d) for d in ds], precision=native.QMS_MIXED, max_active=3,
                          on_result=lambda req, s: got.__setitem__(req.design.name, s.get_fields()))
        assert all(r.error is None for r in res.records), [r.error for r in res.records]
        assert res.peak_active > 1
        for d in ds:
            assert np.array_equal(got[d.name], alone[d.name]), d.name
        d = ds[0]
        with serve.make_solver(d, native.QMS_MIXED) as s:
            s.steady(serve.program.steady_velocity(d.inputs.traj, d.inputs.dt))
            s.run_async(serve.velocities(d, 1, d.inputs.steps))
            try:
                s.get_fields()
                raise AssertionError("busy handle answered")
            except native.QmsError:
                pass
            s.wait()
            assert np.array_equal(s.get_fields(), alone[d.name])
        print("ok")
    """)
    assert "ok" in out
