"""Benchmark of the SCFT hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Headline (`value`): the FP32-propagator path on a batch of 16 production
systems (separations d = 0.14 ... 0.44) held in one context, in
system-iterations/s; each step advances all 16 systems by one iteration.
Beside it: the single production system with the FP32 propagator and with
the plain FP64 path, their ratio, and |dH/H| between the two; and
(`published_2d`) the rates of two published 2D runs as published (525 x 875
and 625 x 675 grids, N = 60, grafted chains).

A step is one SCFT iteration of the production system (64 x 64 x 128 grid,
box 4 x 4 x 8 Rg, AB diblock N = 100, f_A = 0.5, chi N = 20, C = 50,
kappa N = 50, two square plates at d = 0.30): the 1S semi-implicit field
update (5 forward + 3 inverse 3D FFTs and the k-space update) followed by the
density evaluation (field assembly, smearing, 2 x 99 propagator steps, Q and
the q.q-dagger sum).  Inputs are resident in HBM before timing starts; K
iterations run as replays of one captured HIP graph, bracketed by
hipDeviceSynchronize (scft_synchronize).  torch is not imported: the library
brings its own HIP runtime and is the only GPU client of the process.

`value` is FP32-propagator iterations/s (complex64 propagator and FFTs,
complex128 fields, FP64 Q, log Q and H).  The plain complex128 path is timed
the same way and reported beside it, with |dH/H| between the two after the
same iterations from the same start.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "aligned-nanoplate-separation-and-rotation_amd"))

import scft_amd  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOAD = "production-64x64x128-N100"
METRIC = "SCFT iter/s at 64x64x128 N=100: FP32-batched vs FP64 plain, %HBM roofline"


def b_iter(M, N, b, bf):
    """Algorithmic bytes of one SCFT iteration (reference dataflow: q and
    q-dagger stored, each 3D FFT one read + one write)."""
    return M * (10 * (N - 1) * b + 2 * b + 2 * N * b + 2 * bf + 24 * bf + 12 * bf)


def traffic_for(kernel, systems=1):
    """HBM bytes per launch of a propagator pass from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated on a plain copy),
    per system times the systems of one launch."""
    path = os.path.join(REPO, "profiles", "traffic_64x64x128_N100.json")
    try:
        return systems * json.load(open(path))["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def run_path(p, fields, prec, warmup, steps):
    # torch is not imported: the library is the only HIP client of this process
    ctx = scft_amd.Context(p, device=0, precision=prec)
    ctx.set_fields(*fields)
    ctx.calc_poly_density()
    ctx.iterate(max(warmup, 1))  # includes the eager iteration and graph capture
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.iterate(steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    H = ctx.calc_H()[0]
    return ctx, dt, H


def cpu_baseline(p, fields, budget_s=20.0, threads=None):
    """FP64 oracle (reference dataflow) on the host cores (all of them, or
    `threads`), bounded sample; the set-up density evaluation is not timed."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle

    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(p)
    o.set_fields(*fields)
    o.calc_poly_density()
    n = 0
    t0 = time.perf_counter()
    while True:
        o.update_1s()
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    dt = time.perf_counter() - t0
    o.close()
    return {
        "value": n / dt,
        "unit": "iter/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} 1S iterations of the 64x64x128 N=100 system (FP64 oracle, OpenMP, "
                  f"{dt:.1f} s)",
    }


def run_batch(ws, prec, warmup, steps):
    ps = [w.params() for w in ws]
    ctx = scft_amd.Context(ps, device=0, precision=prec)
    for i, (w, p) in enumerate(zip(ws, ps)):
        ctx.set_fields(*scft_amd.seeded_fields(p, w.seed, period_rg=4.0, amp=1.0, sigma=0.01), sys=i)
    ctx.calc_poly_density()
    ctx.iterate(max(warmup, 1))
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.iterate(steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    return ctx, dt, ps


def published_2d(iters=200):
    """Rates of two published 2D runs as published (525 x 875 and 625 x 675,
    N = 60, grafted chains, the input's own start fields): FP64 and FP32
    iterations/s of one system, with the B_iter fraction of HBM (B_iter of
    the diblock path; the Ng-step graft chains are not counted)."""
    import tempfile
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    from published_input import published
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for run in ("angle-05.00", "dist-0.300"):
            path = os.path.join(td, "bcp.input")
            open(path, "w").write(published(run)[0])
            p = scft_amd.read_input(path, 2)
            for name, prec in (("fp64", scft_amd.PREC_FP64), ("fp32", scft_amd.PREC_FP32)):
                ctx = scft_amd.Context(p, precision=prec)
                ctx.calc_poly_density()
                ctx.iterate(5)
                ctx.synchronize()
                t0 = time.perf_counter()
                ctx.iterate(iters)
                ctx.synchronize()
                rate = iters / (time.perf_counter() - t0)
                ctx.close()
                b = b_iter(p.M, p.N, 16 if name == "fp64" else 8, 16)
                out[f"{run} {p.Nx[0]}x{p.Nx[1]} N={p.N} Ng={p.Ng} {name}"] = {
                    "iter_s": rate, "bytes_per_iter": b, "hbm_frac": b * rate / (HBM_PEAK_GBS * 1e9)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-published", action="store_true", help="skip the published 2D grids")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X "
              "(independent systems would shard across GPUs without exchange; no multi-GPU path).")
        sys.exit(2)

    # headline: FP32 propagator, a batch of 16 production systems in one context
    batch = scft_amd.sweep("separation-64")
    ctxb, dtb, psb = run_batch(batch, scft_amd.PREC_FP32, args.warmup, args.steps)
    nb = len(batch)
    batched = nb * args.steps / dtb
    k_plane = ctxb.time_kernel(0, 20)
    k_z = ctxb.time_kernel(1, 20)
    k_den = ctxb.time_kernel(2, 5)
    dom_name, (dom_ms, dom_bytes) = max((("prop_plane", k_plane), ("prop_z", k_z)),
                                        key=lambda kv: kv[1][0])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    planb = ctxb.describe()
    ctxb.close()
    # the same batch with sqrt(N)-checkpointed propagators (memory-bounded
    # mode: about 3 sqrt(N) instead of 2 N slices per system)
    ctxc, dtc, _ = run_batch(batch, scft_amd.PREC_FP32 | scft_amd.OPT_CHECKPOINT, args.warmup, args.steps)
    batched_ckpt = nb * args.steps / dtc
    ckpt_storage = [l for l in ctxc.describe().splitlines() if l.startswith("propagator_storage")]
    ctxc.close()

    # the single production system, FP32 propagator and plain FP64
    w = scft_amd.config(WORKLOAD)
    p = w.params()
    fields = scft_amd.seeded_fields(p, w.seed, period_rg=4.0, amp=1.0, sigma=0.01)
    M, N = p.M, p.N
    ctx32, dt32, H32 = run_path(p, fields, scft_amd.PREC_FP32, args.warmup, args.steps)
    it32 = args.steps / dt32
    s_plane = ctx32.time_kernel(0, 50)
    s_z = ctx32.time_kernel(1, 50)
    ctx32.time_kernel(4, 20)   # counter-calibration copy for the rocprof runs
    ctx32.close()
    ctx64, dt64, H64 = run_path(p, fields, scft_amd.PREC_FP64, args.warmup, args.steps)
    it64 = args.steps / dt64
    ctx64.close()

    out = {
        "metric": METRIC,
        "value": batched,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dtb / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "batched-16x-production-64x64x128-N100",
            "systems": nb,
            "grid": list(p.grid()),
            "box_rg": [4.0, 4.0, 8.0],
            "N": N,
            "separations": [x.d for x in batch],
            "precision": "complex64 propagator/FFT, complex128 fields, FP64 Q/H",
            "parallelism": "single GPU, systems batched per launch",
        },
        "value_definition": "system-iterations/s of the FP32 batch (16 systems x 1 iteration per step)",
        "single_system_fp32_iter_s": it32,
        "fp32_batched_checkpointed_iter_s": batched_ckpt,
        "checkpointed_storage": ckpt_storage[0] if ckpt_storage else None,
        "fp64_plain_iter_s": it64,
        "fp32_batched_over_fp64_plain": batched / it64,
        "fp32_single_over_fp64_plain": it32 / it64,
        "dH_rel_fp32_vs_fp64": abs(H32 - H64) / abs(H64),
        "H_fp64": [H64.real, H64.imag],
        "iteration_roofline": {
            "bytes_per_iter_fp32": b_iter(M, N, 8, 16),
            "frac_fp32_batched": b_iter(M, N, 8, 16) * batched / (HBM_PEAK_GBS * 1e9),
            "frac_fp32_single": b_iter(M, N, 8, 16) * it32 / (HBM_PEAK_GBS * 1e9),
            "bytes_per_iter_fp64": b_iter(M, N, 16, 16),
            "frac_fp64_single": b_iter(M, N, 16, 16) * it64 / (HBM_PEAK_GBS * 1e9),
        },
        "kernels_ms": {"batch_prop_plane": k_plane[0], "batch_prop_z": k_z[0], "batch_density": k_den[0],
                       "single_prop_plane": s_plane[0], "single_prop_z": s_z[0]},
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name + " (16-system batch, one launch per contour step)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_for(dom_name, nb),
        },
        "cpu_baseline": None,
    }
    if not args.no_published:
        out["published_2d"] = published_2d()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, fields, args.cpu_budget)
        # the same oracle on one core (at least one iteration)
        out["cpu_baseline_1core"] = cpu_baseline(p, fields, args.cpu_budget / 2, threads=1)
    print(planb, file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
