#!/usr/bin/env python3
"""Benchmark of the QMS_3D time-stepping core on one MI355X.

Workload (the configuration the headline "A-V time steps/s at 128^3" is
quoted on): config 3, 128^3 nodes (8,339,456 DoF), 0.25 mm cells, 24 random
boxes plus a 40x40x8-cell conducting slab, frame velocity ramped 0 -> 5 m/s
over 200 steps of 2 us (the reference's constant-acceleration trajectory,
QMS_3D.f90:2331-2350; with a constant velocity every step reproduces the
steady state, so the ramp is the variant whose steps carry work).

Per step (QMS_3D.f90:1424-2234 restated on the GPU): right-hand side,
Dirichlet values, extrapolated initial guess, FGMRES to a scaled relative
residual of 3e-13 with the block-triangular multigrid preconditioner, every
iteration a replayed hipGraph; a step whose FGMRES stagnates at the residual's
rounding floor (which rises to about 1e-11 as the fields grow over the run)
is accepted within 1e-10 and counted (`stalled_solves`).  3e-13 is what keeps
the mixed path's A, B and sigma*grad(V) within 1e-8 of the FP64 path's at
every checked step (at 1e-12 sigma*grad(V) differs by 2e-8 at step 13, where
both solves still converge).  Fields stay on the device between steps.

Timed: the steady state, `--warmup` untimed steps, then (by default) every
remaining step of the 200-step run, each a blocking call; `ms_per_step` is
the mean of the timed steps.  The frame velocity grows over the run, and so
do the FGMRES iterations per step: the early steps alone would overstate the
rate.  The
mixed-precision path (FP64 refinement around an inner FGMRES with the A
components of its Krylov vectors, its Z basis and its A-block multigrid in
FP32; V components, V-block multigrid, residuals and fields FP64) is
`value`; the plain FP64 path runs the same steps beside it and every timed
step's A, B and sigma*grad(V) are compared with it.

`roofline` is the dominant kernel of the mixed step, the level-0 A-block
smoother sweep, timed with HIP events on the handle's stream; `kernels` holds
the other kernels of the step at 128^3 (operator figures with cold caches:
each launch reads copies of its arrays that the other launches' traffic has
evicted).  `batch` applies config 4's 64 operators in one launch;
`batch_steps` steps config 4's designs together; `serving` serves a config-5
request stream and reports p50/p99 latency.  `cpu_baseline` is the oracle (the C restatement of
the reference's CSR assembly with SciPy's SuperLU in place of PARDISO) on one
core: config 1's design at 24^3 nodes (its 32^3 factorisation takes over two
minutes on one core), steady factorisation then steps that reuse it, as the
reference's phase 33 does for a constant velocity; the GPU runs the same
design and steps and its errors against the port are reported there.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nodes N] [--paths both|mixed|fp64]
                       [--no-kernels] [--no-graphs] [--no-cpu-baseline] [--no-batch] [--no-serve]
"""
import argparse
import json
import os
import shutil
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TOL = 1e-12
# the headline's solves: the tolerance that keeps mixed within 1e-8 of FP64 at
# every checked step, and the residual a stagnating solve is accepted at
HEADLINE_TOL = 3e-13
HEADLINE_FLOOR_TOL = 1e-10


def traffic_from_profile(name: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary profiles/traffic_<name>.txt (FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by what the counter reads for a plain copy kernel of
    known size), or (None, None) when no such summary is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.txt")
    if not os.path.exists(path):
        return None, None
    kv = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        t = float(kv["fetch_kib"]) * float(kv["fetch_scale"]) + float(kv["write_kib"]) * float(kv["write_scale"])
    except (KeyError, ValueError):
        return None, None
    return t * 1024.0, os.path.relpath(path, ROOT)


def make_solver(d, precision, **kw):
    from alpineqms_amd import native
    from alpineqms_amd.qms_io import material_slots
    inp = d.inputs
    slots, props = material_slots(inp, d.geom)
    return native.Solver(inp.cells, inp.h, inp.dt, inp.bc_vector(), slots, props, precision=precision,
                         tol=kw.pop("tol", TOL), **kw)


def state_errors(d, coef, f, fr, B, Br):
    from alpineqms_amd.fields import a_part, rel, sigma_grad_v
    h = d.inputs.h
    return {"A": rel(a_part(f), a_part(fr)), "B": rel(B, Br),
            "sigma_grad_V": rel(sigma_grad_v(h, coef, f), sigma_grad_v(h, coef, fr))}


def reference_binary_probe():
    """The reference binary needs Intel ifx (or another Fortran compiler) and
    oneMKL's PARDISO (3D/compile.sh:4-12); report what this host has."""
    import ctypes.util
    found = {c: shutil.which(c) for c in ("ifx", "ifort", "gfortran", "flang")}
    mkl = ctypes.util.find_library("mkl_rt")
    ok = bool(mkl) and any(found[c] for c in ("ifx", "ifort", "gfortran"))
    return {"available": ok, "compilers": {k: v for k, v in found.items() if v}, "libmkl_rt": mkl,
            "note": None if ok else "no Fortran compiler with oneMKL PARDISO on this host: the reference binary "
                                    "cannot be built, so the baseline is the oracle port"}


def cpu_baseline(nsteps: int = 10):
    """The oracle port on one core, config 1's design at 24^3 nodes, and the
    GPU paths on the same design and steps checked against it."""
    import numpy as np
    from oracle import oracle as O
    from alpineqms_amd import configs, native, program
    from alpineqms_amd.fields import a_part, rel
    from tests.helpers import problem
    d = configs.config1(24, steps=nsteps)
    p = problem(d)
    coef = O.coeffs(p)
    t0 = time.perf_counter()
    M, b = O.assemble(p, coef, O.steady_velocity(p))
    S = O.DirectSolver(M)
    f = O.unpack(p, S.solve(b))
    t_factor = time.perf_counter() - t0
    ref = [f]
    # constant frame velocity: the factorisation is reused, as phase 33 does
    v = O.step_velocity(p, 1)
    M, b = O.assemble(p, coef, v, p.dt, f)
    S = O.DirectSolver(M)
    t1 = time.perf_counter()
    for n in range(1, nsteps + 1):
        M, b = O.assemble(p, coef, O.step_velocity(p, n), p.dt, f)    # per-step CSR rebuild (:1452-2191)
        S.M = M.tocsr()
        f = O.unpack(p, S.solve(b))                                   # triangular solves + 2 refinement steps
        ref.append(f)
    t_step = (time.perf_counter() - t1) / nsteps
    out = {"value": 1.0 / t_step, "unit": "steps/s", "cores": 1, "kind": "port",
           "sample": f"config 1 design at 24^3 nodes ({d.dof} DoF; 31^3 cells scaled to 23^3, 12^3 block), "
                     f"steady factorisation {t_factor:.1f} s, then {nsteps} steps of CSR assembly + "
                     f"triangular solves with the step-1 factorisation ({t_step:.3f} s per step)",
           "dof_steps_per_s": d.dof / t_step, "factorisation_s": t_factor,
           "reference_binary": reference_binary_probe()}
    errs = {}
    for name, prec in (("fp64", native.QMS_FP64), ("mixed", native.QMS_MIXED)):
        with make_solver(d, prec, tol=1e-12) as s:
            s.steady(program.steady_velocity(d.inputs.traj, d.inputs.dt))
            e = []
            for n in range(1, nsteps + 1):
                s.step(program.step_velocity(d.inputs.traj, d.inputs.dt, n))
                g = s.get_fields()
                e.append(max(rel(a_part(g), a_part(ref[n])),
                             rel(O.bfield(p, g), O.bfield(p, ref[n]))))
            errs[name] = max(e)
    out["gpu_max_rel_error_vs_port_A_B"] = errs
    return out


def smoother_roofline(s, mixed: bool, reps: int):
    from alpineqms_amd import native
    s.time_kernel(native.K_SMOOTH_A, 3)
    ms, nb = s.time_kernel(native.K_SMOOTH_A, reps)
    gbs = nb / (ms * 1e-3) / 1e9
    return ms, nb, gbs


def batch_config4(reps: int, ndesigns: int = 64):
    """Config 4's 64 designs: the operator of every design in one persistent
    batched launch against one launch per design."""
    import numpy as np
    from alpineqms_amd import configs, native
    ds = configs.config4_designs(ndesigns)
    rng = np.random.default_rng(4)
    ss = []
    for d in ds:
        s = make_solver(d, native.QMS_FP64, restart=1)
        nx, ny, nz = s.nodes
        s.set_fields(rng.normal(size=(4, nz, ny, nx)))
        ss.append(s)
    v = rng.normal(size=(len(ss), 3))
    dof = sum(d.dof for d in ds)
    with native.Batch(ss) as b:
        seq0, _ = b.time(v, reps=reps, sequential=True)
        kz, cw, blocks = b.tune(v, reps=10)
        ms_b, nbytes = b.time(v, reps=reps)
    for s in ss:
        s.close()
    gbs = nbytes / (ms_b * 1e-3) / 1e9
    return {"workload": f"config 4: {len(ds)} designs, 24-96 nodes per axis, {dof} DoF, transient operator rows",
            "dof": dof, "batched_ms": ms_b, "sequential_default_ms": seq0, "achieved_GBs": gbs,
            "frac": gbs / HBM_PEAK_GBS, "speedup_vs_sequential": seq0 / ms_b, "blocks": blocks}


def serving_config5(count: int, rate: float, max_active: int, ramp_v: float = 0.0, smax: int = 200,
                    tol: float = TOL):
    """Config 5 as a served request stream: Poisson arrivals, cubic designs of
    32-160 nodes per axis, 20-smax steps each, admitted by device footprint
    under 288 GB with at most `max_active` in flight, mixed precision.
    ramp_v != 0: the dynamic variant (frame velocity ramped 0 -> ramp_v m/s
    over each request's steps, so every step carries a solve)."""
    from alpineqms_amd import configs, native, serve
    reqs = configs.config5_stream(count, rate, smax=smax, ramp_v=ramp_v)
    res = serve.serve(reqs, precision=native.QMS_MIXED, max_active=max_active, tol=tol)
    ok = [r for r in res.records if r.error is None]
    vdesc = (f"frame velocity ramped 0 -> {ramp_v:g} m/s along x over each request's steps (the dynamic variant)"
             if ramp_v else "frame velocity 0 (the literal config: each step after the steady state reproduces it)")
    return {"workload": f"config 5: {count} requests, Poisson arrivals at {rate:g}/s (rng 5000), N^3 with N "
                        f"log-uniform over multiples of 8 in [32, 160], 20-{smax} steps of 5 us each, {vdesc}, "
                        f"mixed precision to {tol:g} scaled residual, at most {max_active} in flight, 288 GB budget",
            "requests": count, "completed": len(ok), "errors": [r.error for r in res.records if r.error],
            "p50_latency_s": res.percentile(50), "p99_latency_s": res.percentile(99),
            "mean_wait_s": sum(r.wait for r in ok) / max(1, len(ok)),
            "dof_steps_per_s": res.dof_steps_per_s, "wall_s": res.wall_s,
            "peak_active": res.peak_active, "peak_device_bytes": res.peak_bytes,
            "sizes": [r.name.split("_")[-1] for r in res.records],
            "iterations": [r.iters for r in res.records]}


def batch_steps_config4(ndesigns: int, steps: int, ramp_v: float = 0.0, tol: float = TOL):
    """Config 4's designs stepped together (each on its own stream from a
    library thread) against the same steps one design at a time."""
    from alpineqms_amd import configs, native, serve
    ds = configs.config4_designs(ndesigns, steps=steps, ramp_v=ramp_v)
    r = serve.run_batch(ds, native.QMS_MIXED, steps, tol=tol)
    vdesc = (f"frame velocity ramped 0 -> {ramp_v:g} m/s along x (the dynamic variant)" if ramp_v
             else "frame velocity 0 (the literal config)")
    return {"workload": f"config 4: {ndesigns} designs, 24-96 nodes per axis, {sum(d.dof for d in ds)} DoF, "
                        f"{steps} steps each after the steady state, {vdesc}, mixed precision to {tol:g} scaled "
                        f"residual",
            "together_s": r.batched_s, "one_at_a_time_s": r.sequential_s, "speedup": r.speedup,
            "dof_steps_per_s": r.dof_steps / r.batched_s, "iterations": r.iters}


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps after the warmup (0: every remaining step of the 200-step run)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--check-every", type=int, default=10,
                    help="compare the paths' A, B and sigma*grad(V) every this many timed steps (and at the last)")
    ap.add_argument("--nodes", type=int, default=128)
    ap.add_argument("--tol", type=float, default=HEADLINE_TOL, help="scaled relative residual of every solve")
    ap.add_argument("--floor-tol", type=float, default=HEADLINE_FLOOR_TOL,
                    help="scaled residual at which a stagnating solve is accepted")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch each FGMRES iteration's kernels one by one instead of replaying captured graphs "
                         "(the same kernels and results; for profilers that do not follow graph launches)")
    ap.add_argument("--paths", choices=("both", "mixed", "fp64"), default="both",
                    help="which precision paths to step (both: mixed is value, FP64 beside it)")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel timings")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    return ap


def main():
    a = parser().parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: every design of the benchmark set fits one MI355X, and the "
              "multi-GPU path (whole designs per GPU) is not built", file=sys.stderr)
        return 2
    from alpineqms_amd import configs, native, program
    d = configs.config3(N=a.nodes, ramp=True)
    inp = d.inputs
    nx, ny, nz = d.nodes
    print(f"bench: {d.name} nodes={d.nodes} dof={d.dof}", file=sys.stderr, flush=True)
    names = {"both": ("fp64", "mixed"), "mixed": ("mixed",), "fp64": ("fp64",)}[a.paths]
    precs = {"fp64": native.QMS_FP64, "mixed": native.QMS_MIXED}
    runs = {}
    solvers = {}
    for name in names:
        s = make_solver(d, precs[name], tol=a.tol, floor_tol=a.floor_tol, use_graphs=not a.no_graphs)
        t0 = time.perf_counter()
        s.steady(program.steady_velocity(inp.traj, inp.dt))
        st = s.stats()
        print(f"bench: {name} steady state {time.perf_counter() - t0:.2f} s, {st.last_iters} iterations",
              file=sys.stderr, flush=True)
        solvers[name] = s
        runs[name] = {"steady_iters": st.last_iters, "step_ms": [], "iters": []}
    coef = next(iter(solvers.values())).get_coeffs()
    nsteps = a.steps if a.steps > 0 else inp.steps - a.warmup
    if a.warmup + nsteps > inp.steps:
        print(f"bench: the run has {inp.steps} steps", file=sys.stderr)
        return 2
    a.steps = nsteps
    errors = {"A": 0.0, "B": 0.0, "sigma_grad_V": 0.0} if len(names) == 2 else None
    checked = []
    err_steps = []
    for n in range(1, a.warmup + a.steps + 1):
        v = program.step_velocity(inp.traj, inp.dt, n)
        for name, s in solvers.items():
            q0 = s.stats().total_iters
            t0 = time.perf_counter()
            s.step(v)
            t = time.perf_counter() - t0
            if n > a.warmup:
                runs[name]["step_ms"].append(t * 1e3)
                runs[name]["iters"].append(int(s.stats().total_iters - q0))
        if n > a.warmup:
            msg = ", ".join(f"{k} {runs[k]['step_ms'][-1]:.2f} ms ({runs[k]['iters'][-1]} it)" for k in names)
            last = n == a.warmup + a.steps
            if errors is not None and ((n - a.warmup) % a.check_every == 0 or last):
                checked.append(n)
                f64, fmx = solvers["fp64"].get_fields(), solvers["mixed"].get_fields()
                B64, Bmx = solvers["fp64"].get_h(unit_mu=True), solvers["mixed"].get_h(unit_mu=True)
                e = state_errors(d, coef, fmx, f64, Bmx, B64)
                for k in errors:
                    errors[k] = max(errors[k], e[k])
                err_steps.append({"step": n, **e})
                msg += f", mixed vs fp64: A {e['A']:.1e} B {e['B']:.1e} sgV {e['sigma_grad_V']:.1e}"
            print(f"bench: step {n}: {msg}", file=sys.stderr, flush=True)
    for name in runs:
        r = runs[name]
        r["stalled_solves"] = int(solvers[name].stats().stalled_solves)
        r["device_bytes"] = int(solvers[name].stats().device_bytes)
        r["total_s"] = sum(r["step_ms"]) / 1e3
        r["ms_per_step"] = sum(r["step_ms"]) / len(r["step_ms"])
        r["steps_per_s"] = 1e3 / r["ms_per_step"]
        r["dof_steps_per_s"] = d.dof * r["steps_per_s"]
        r["mean_iters"] = sum(r["iters"]) / len(r["iters"])
        del r["step_ms"]
    # the dominant kernel of the step: the level-0 A-block smoother sweep (FP32 on the mixed path)
    main_name = "mixed" if "mixed" in solvers else "fp64"
    ms_s, nb_s, gbs_s = smoother_roofline(solvers[main_name], main_name == "mixed", 50)
    kernels = None
    if not a.no_kernels:
        kernels = {}
        vlast = program.step_velocity(inp.traj, inp.dt, a.warmup + a.steps)
        s64 = solvers.get("fp64")
        if s64 is not None:
            ms_s64, nb_s64, gbs_s64 = smoother_roofline(s64, False, 50)
            kernels["smooth_A_fp64"] = {"ms_per_launch": ms_s64, "algorithmic_bytes": nb_s64,
                                        "achieved_GBs": gbs_s64, "frac": gbs_s64 / HBM_PEAK_GBS}
            for kname, which in (("operator_apply", native.K_APPLY), ("operator_resid", native.K_RESID),
                                 ("rhs", native.K_RHS), ("hfield", native.K_HFIELD), ("copy", native.K_COPY),
                                 ("smooth_V_fp64", native.K_SMOOTH_V), ("gram_schmidt_31", native.K_GRAM_SCHMIDT)):
                sel = which | (native.K_COLD if which in (native.K_APPLY, native.K_RESID, native.K_RHS,
                                                           native.K_HFIELD, native.K_COPY) else 0)
                s64.time_kernel(sel, 2, v=vlast)
                ms, nb = s64.time_kernel(sel, 20, v=vlast)
                gbs = nb / (ms * 1e-3) / 1e9
                kernels[kname] = {"ms_per_launch": ms, "algorithmic_bytes": nb, "achieved_GBs": gbs,
                                  "frac": gbs / HBM_PEAK_GBS, "cold_caches": bool(sel & native.K_COLD)}
            # the same copy kernel on 2 GiB arrays: the sustained rate of a plain stream
            s64.time_kernel(native.K_COPY_SUSTAINED, 1)
            ms, nb = s64.time_kernel(native.K_COPY_SUSTAINED, 10)
            gbs = nb / (ms * 1e-3) / 1e9
            kernels["copy_sustained_4GiB"] = {"ms_per_launch": ms, "algorithmic_bytes": nb, "achieved_GBs": gbs,
                                              "frac": gbs / HBM_PEAK_GBS, "cold_caches": True}
    for s in solvers.values():
        s.close()
    head = runs[main_name]
    traffic_s, traffic_src = traffic_from_profile(f"k_a0_zm_{'fp32' if main_name == 'mixed' else 'fp64'}_sweep_"
                                                  f"config3_{a.nodes}")
    sw = "FP32" if main_name == "mixed" else "FP64"
    ssz = 4 if main_name == "mixed" else 8
    line = {
        "metric": "A-V time steps/s at 128^3, mixed vs FP64 @1e-8 rel err; served p50/p99 latency",
        "value": head["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_name == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"config 3: {nx}x{ny}x{nz} nodes, 0.25 mm, 24 boxes + 40x40x8 slab, frame velocity "
                               f"ramp 0 -> 5 m/s over 200 steps of 2 us, steps {a.warmup + 1}-"
                               f"{a.warmup + a.steps} timed, {main_name} path, FGMRES(30) to {a.tol:g} scaled "
                               f"residual (accepted at stagnation within {a.floor_tol:g})",
                   "dof": d.dof},
        "paths": dict(runs, **({"mixed_over_fp64": runs["mixed"]["steps_per_s"] / runs["fp64"]["steps_per_s"]}
                               if len(names) == 2 else {})),
        "dof_steps_per_s": head["dof_steps_per_s"],
        "max_rel_error_mixed_vs_fp64": errors,
        "rel_error_mixed_vs_fp64_by_step": err_steps,
        "roofline": {"kernel": f"k_a0_zm (level-0 A-block smoother sweep (weighted Jacobi), {sw}, 3 components per node, "
                               f"LDS z-march, coefficients "
                               f"{'from the six FP32 arrays' if main_name == 'mixed' else 'through the class table'})",
                     "bound": "hbm", "achieved": gbs_s, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs_s / HBM_PEAK_GBS, "traffic": traffic_s, "traffic_source": traffic_src,
                     "algorithmic_bytes": nb_s,
                     "algorithmic_bytes_definition": f"{ssz}*(9*N^3 + 6*N^3): x, b and x_out (3 components each) "
                                                     f"and the six {sw} coefficient arrays",
                     "ms_per_launch": ms_s, "launches": 50,
                     "caches": "back to back on the handle's arrays: the sweeps of a V-cycle reuse the same b and "
                               "coefficients, so this is how they run inside a step"},
        "kernels": kernels,
        "batch": None if a.no_batch else batch_config4(20),
        # the ramp variants at 1e-10: on these 0.5 mm designs restarted FGMRES(30) can stagnate between
        # 1e-12 and 1e-11 (the conductors' floating potentials), which the headline's 0.25 mm design does not
        "batch_steps": None if a.no_batch else {"literal": batch_steps_config4(64, 10),
                                                "ramp": batch_steps_config4(64, 10, ramp_v=1.0, tol=1e-10)},
        "serving": None if a.no_serve else {"literal": serving_config5(24, 4.0, 8),
                                            "ramp": serving_config5(24, 4.0, 8, ramp_v=1.0, smax=60, tol=1e-10)},
        "cpu_baseline": None if a.no_cpu_baseline else cpu_baseline(),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
