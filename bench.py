#!/usr/bin/env python3
"""Benchmark of the QMS_3D time-stepping core on one MI355X.

Workload: the 128^3-node design of the benchmark set (config 3: 0.25 mm
cells, 24 random boxes plus a 40x40x8-cell conducting slab, frame velocity
ramped to 5 m/s over 200 steps of 2 us), the configuration the headline
metric "A-V time steps/s at 128^3" is quoted on.

Printed: one JSON line.  The time-step rate needs the multigrid solve path,
which is disabled in the library for now (it faulted the GPU during its first
hierarchy build, cause not identified), so `value` is null and says so.
What is measured on the GPU: the fused A-V operator kernel at 128^3 (the
matrix-free replacement of the reference's per-step CSR assembly and the
operator inside every Krylov iteration) with the rows of the last ramp step
(transient, v = 5 m/s), against the HBM roofline, with HIP events on the
stream it is launched on; its column shape and z-chunk depth (planes per
workgroup) are picked from a sweep.  Beside it: the FP64 residual, the step right-hand side, the
cell-centred H, a plain copy kernel of the same state (calibration) and
the FGMRES Gram-Schmidt step against 31 basis vectors (multi-dot,
reduction, update, norm).
Every kernel figure is timed with cold caches, as inside a solve, where the
Krylov basis streams through the caches between two launches of a kernel:
back-to-back launches rotate through copies of the kernel's arrays, enough
that 640 MB of other arrays (more than the L2s and the 256 MB Infinity Cache
hold) pass between two uses of one copy.  Back-to-back launches (reported beside,
as `warm_repeated`) can keep a working set below 256 MB, such as the apply's
142 MB at 128^3, in the Infinity Cache and run faster than HBM allows.
The `batch` object measures the config-4 workload (64 uneven designs): the
operator of every design in one persistent batched launch against one
launch per design.  The CPU baseline is the oracle (CSR assembly +
equilibrated SuperLU factor/solve, the restated reference algorithm) timed
on one core over ramp steps 1-3 of a 24^3-node design drawn from the same
box distribution.  `operator_sizes` repeats the apply at 32^3, 64^3, 96^3 and
160^3 nodes, each at its fastest column shape.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--zchunk KZ --colwidth CW] [--no-cpu-baseline]
       [--no-batch] [--no-sizes] [--only-batch]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_baseline():
    import numpy as np
    from oracle import oracle as O
    from alpineqms_amd import configs
    from tests.helpers import problem
    d = configs.config2(N=24, steps=10, ramp_v=5.0)
    p = problem(d)
    coef = O.coeffs(p)
    M, b = O.assemble(p, coef, O.steady_velocity(p))
    f = O.unpack(p, O.DirectSolver(M).solve(b))
    nsteps = 3
    t0 = time.perf_counter()
    for n in range(1, nsteps + 1):
        v = O.step_velocity(p, n)
        M, b = O.assemble(p, coef, v, p.dt, f)      # per-step CSR rebuild (QMS_3D.f90:1452-2191)
        x = O.DirectSolver(M).solve(b)              # factor + solve (phase 23: the velocity changed)
        f = O.unpack(p, x)
    t = (time.perf_counter() - t0) / nsteps
    return {"value": 1.0 / t, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"ramp steps 1-3 of a 24^3-node design ({d.dof} DoF, 8 boxes, v ramp to 5 m/s): "
                      f"CSR assembly + row-equilibrated SuperLU factor and solve, {t:.2f} s per step",
            "dof_steps_per_s": d.dof / t}


SHAPES = ((64, 4), (64, 6), (64, 8), (32, 4), (32, 6), (32, 8), (32, 12), (32, 16))   # (column width, z-chunk)


def traffic_from_profile(kernel: str, shape: str):
    """HBM bytes per launch of `kernel` at column shape `shape` ("32x8": column
    width x z-chunk) from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by what the counter reads for the
    plain copy kernel of known size), or None when no summary of that shape
    is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{shape}.txt")
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


def batch_config4(reps: int, ndesigns: int = 64):
    """Config 4: 64 independent designs (24..96 nodes per axis, 4-12 boxes
    each), operator rows of a transient step at a random frame velocity per
    design, applied in one persistent batched launch (z-chunks and grid size
    from the autotuner) and, for comparison, as one launch per design."""
    import numpy as np
    from alpineqms_amd import configs, native
    from alpineqms_amd.qms_io import material_slots
    ds = configs.config4_designs(ndesigns)
    rng = np.random.default_rng(4)
    ss = []
    for d in ds:
        slots, props = material_slots(d.inputs, d.geom)
        s = native.Solver(d.inputs.cells, d.inputs.h, d.inputs.dt, d.inputs.bc_vector(), slots, props, restart=1)
        nx, ny, nz = s.nodes
        s.set_fields(rng.normal(size=(4, nz, ny, nx)))
        ss.append(s)
    v = rng.normal(size=(len(ss), 3))
    dof = sum(d.dof for d in ds)
    print(f"bench: config 4 batch of {len(ds)} designs, {dof} DoF", file=sys.stderr, flush=True)
    with native.Batch(ss) as b:
        seq0, _ = b.time(v, reps=reps, sequential=True)          # default shape, one launch per design
        # best shape of each design on its own, then one launch per design
        for s_, vd in zip(ss, v):
            best = None
            for cw_, kz_ in SHAPES + ((64, 1), (64, 2), (32, 1), (32, 2)):
                s_.set_colwidth(cw_)
                s_.set_zchunk(kz_)
                s_.time_kernel(native.K_APPLY, 2, v=vd)
                t = s_.time_kernel(native.K_APPLY, 5, v=vd)[0]
                if best is None or t < best[0]:
                    best = (t, cw_, kz_)
            s_.set_colwidth(best[1])
            s_.set_zchunk(best[2])
        ms_s, _ = b.time(v, reps=reps, sequential=True)
        kz, cw, blocks = b.tune(v, reps=10)
        ms_b, nbytes = b.time(v, reps=reps)
    for s in ss:
        s.close()
    gbs = nbytes / (ms_b * 1e-3) / 1e9
    return {"workload": f"config 4: {len(ds)} designs, 24-96 nodes per axis, {dof} DoF, transient rows",
            "dof": dof, "algorithmic_bytes": nbytes, "zchunks": kz, "colwidths": cw, "blocks": blocks,
            "batched_ms": ms_b, "sequential_tuned_ms": ms_s, "sequential_default_ms": seq0,
            "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS, "dof_applies_per_s": dof / (ms_b * 1e-3),
            "speedup_vs_sequential": min(ms_s, seq0) / ms_b}


def size_scan(reps: int, sizes=(32, 64, 96, 160)):
    """The same apply (config-3 box layout and v = 5 m/s rows, spacing scaled
    to the size) at other node counts per axis: config 1/2 sizes and the
    largest served size, each at the fastest column shape of the sweep (a
    small design alone cannot fill the GPU: 32^3 nodes are 128 columns of
    32 x 8 nodes even one plane deep).  Algorithmic bytes and timing (cold
    caches, warm beside) as for the headline."""
    import numpy as np
    from alpineqms_amd import configs, native, program
    from alpineqms_amd.qms_io import material_slots
    out = {}
    for n in sizes:
        d = configs.config3(N=n, ramp=True)
        inp = d.inputs
        slots, props = material_slots(inp, d.geom)
        s = native.Solver(inp.cells, inp.h, inp.dt, inp.bc_vector(), slots, props, restart=1)
        nx, ny, nz = d.nodes
        s.set_fields(np.random.default_rng(n).normal(size=(4, nz, ny, nx)))
        v = program.step_velocity(inp.traj, inp.dt, inp.steps)
        best = None
        for cw, kz in SHAPES + ((64, 1), (64, 2), (32, 1), (32, 2)):
            s.set_colwidth(cw)
            s.set_zchunk(kz)
            s.time_kernel(native.K_APPLY, 2, v=v)
            ms = s.time_kernel(native.K_APPLY | native.K_COLD, 5, v=v)[0]
            if best is None or ms < best[0]:
                best = (ms, cw, kz)
        s.set_colwidth(best[1])
        s.set_zchunk(best[2])
        s.time_kernel(native.K_APPLY, 3, v=v)
        ms_warm, nbytes = s.time_kernel(native.K_APPLY, reps, v=v)
        ms, _ = s.time_kernel(native.K_APPLY | native.K_COLD, reps, v=v)
        s.close()
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[str(n)] = {"dof": d.dof, "column": best[1], "zchunk": best[2], "ms_per_launch": ms,
                       "algorithmic_bytes": nbytes, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
                       "dof_applies_per_s": d.dof / (ms * 1e-3), "warm_ms_per_launch": ms_warm,
                       "warm_frac": nbytes / (ms_warm * 1e-3) / 1e9 / HBM_PEAK_GBS}
        print(f"bench: apply at {n}^3 nodes (column {best[1]}, z-chunk {best[2]}): {ms * 1e3:.1f} us cold "
              f"({gbs / HBM_PEAK_GBS:.2f} of HBM peak), {ms_warm * 1e3:.1f} us warm", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nodes", type=int, default=128)
    ap.add_argument("--zchunk", type=int, default=0,
                    help="planes per operator workgroup (0: sweep column shapes and z-chunks)")
    ap.add_argument("--colwidth", type=int, default=32, help="operator column width with --zchunk (32 or 64)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the config-4 batched operator measurement")
    ap.add_argument("--only-batch", action="store_true", help="print only the config-4 batched measurement")
    ap.add_argument("--no-sizes", action="store_true", help="skip the apply at other node counts")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: every design here fits one MI355X, and the multi-GPU "
              "path (whole designs per GPU, no inter-GPU traffic) is not implemented", file=sys.stderr)
        return 2
    import numpy as np
    from alpineqms_amd import configs, native, program
    from alpineqms_amd.qms_io import material_slots

    if a.only_batch:
        print(json.dumps({"batch": batch_config4(a.steps)}), flush=True)
        return 0
    d = configs.config3(N=a.nodes, ramp=True)
    inp = d.inputs
    slots, props = material_slots(inp, d.geom)
    print(f"bench: {d.name} nodes={d.nodes} dof={d.dof}", file=sys.stderr, flush=True)
    s = native.Solver(inp.cells, inp.h, inp.dt, inp.bc_vector(), slots, props, precision=native.QMS_FP64)
    nx, ny, nz = d.nodes
    rng = np.random.default_rng(0)
    s.set_fields(rng.normal(size=(4, nz, ny, nx)))
    # rows of the last ramp step: transient, frame velocity 5 m/s (advection terms on)
    v = program.step_velocity(inp.traj, inp.dt, inp.steps)
    w = max(a.warmup, 1)

    def timed(which):
        """(cold ms, bytes, warm ms) of one kernel: cold caches before every
        launch, then back-to-back launches."""
        s.time_kernel(which, w, v=v)
        warm, nb = s.time_kernel(which, a.steps, v=v)
        cold, _ = s.time_kernel(which | native.K_COLD, a.steps, v=v)
        return cold, nb, warm

    def rate(ms, nb):
        gbs = nb / (ms * 1e-3) / 1e9
        return {"ms_per_launch": ms, "algorithmic_bytes": nb, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}

    sweep = {}
    shapes = ((a.colwidth, a.zchunk),) if a.zchunk else SHAPES
    for cw, kz in shapes:
        s.set_colwidth(cw)
        s.set_zchunk(kz)
        sweep[(cw, kz)] = timed(native.K_APPLY)
        print(f"bench: operator column {cw}, z-chunk {kz}: {sweep[(cw, kz)][0] * 1e3:.1f} us cold, "
              f"{sweep[(cw, kz)][2] * 1e3:.1f} us warm", file=sys.stderr, flush=True)
    cw, kz = min(sweep, key=lambda k: sweep[k][0])
    s.set_colwidth(cw)
    s.set_zchunk(kz)
    ms, nbytes, ms_warm = timed(native.K_APPLY)
    achieved = nbytes / (ms * 1e-3) / 1e9
    ncls = s.stats().coef_classes
    # the same apply with the six coefficient arrays streamed from HBM
    s.use_coef_classes(False)
    ms_st, bytes_st, ms_st_warm = timed(native.K_APPLY)
    s.use_coef_classes(True)
    others = {}
    for name, which in (("k_operator_resid", native.K_RESID), ("k_rhs", native.K_RHS),
                        ("k_hfield", native.K_HFIELD), ("k_copy", native.K_COPY),
                        ("gram_schmidt_31", native.K_GRAM_SCHMIDT)):
        mk, bk, mw = timed(which)
        others[name] = dict(rate(mk, bk), warm_repeated=rate(mw, bk))
    status = None
    try:
        s.steady(np.zeros(3))
    except native.QmsError as e:
        status = str(e)
    s.close()
    traffic, tsrc = traffic_from_profile("k_operator_apply_config3_128" + ("_classes" if ncls else ""), f"{cw}x{kz}")
    line = {
        "metric": "A-V time steps/s at 128^3, mixed vs FP64 @1e-8 rel err; served p50/p99 latency",
        "value": None,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"config 3: {nx}x{ny}x{nz} nodes, 0.25 mm, 24 boxes + slab, v ramp 0->5 m/s "
                               f"(rows of the last ramp step: transient, v = {v[0]:g} m/s)",
                   "dof": d.dof},
        "status": "time-step rate not measured: " + (status or "solve path unavailable"),
        "roofline": {"kernel": f"k_operator (fused A-V operator apply, FP64, LDS z-march, column {cw}, z-chunk {kz}"
                               f"{', coefficient classes' if ncls else ''})",
                     "bound": "hbm",
                     "algorithmic_bytes_definition": "8*(2n + 6*N^3): x and y (n = 4N^3 doubles each) and the six "
                                                     "FP64 coefficient arrays of the reference operator",
                     "kernel_stream_bytes": (16.0 * 4 * nx * ny * nz + 2.0 * nx * ny * nz + 48.0 * ncls) if ncls
                     else nbytes,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": tsrc,
                     "algorithmic_bytes": nbytes, "ms_per_launch": ms, "launches": a.steps,
                     "caches": "cold: back-to-back launches rotate through copies of x and y, 640 MB of "
                               "other arrays between two uses of one copy (L2s + 256 MB Infinity Cache evicted)",
                     "warm_repeated": dict(rate(ms_warm, nbytes), launches=a.steps)},
        "shape_sweep_ms": {f"{c}x{z}": m[0] for (c, z), m in sweep.items()},
        "shape_sweep_warm_ms": {f"{c}x{z}": m[2] for (c, z), m in sweep.items()},
        "coefficient_classes": ncls,
        "operator_streamed_coefficients": dict(rate(ms_st, bytes_st), warm_repeated=rate(ms_st_warm, bytes_st)),
        "kernels": others,
        "operator_sizes": None if a.no_sizes else size_scan(a.steps),
        "batch": None if a.no_batch else batch_config4(a.steps),
        "cpu_baseline": None if a.no_cpu_baseline else cpu_baseline(),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
