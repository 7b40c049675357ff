"""Errors of H, Q and the densities per benchmark configuration.

Each test runs one configuration on the GPU, compares with the FP64 oracle
(short runs at the full sizes) or with the plain FP64 GPU path (long runs),
asserts the bounds and prints one line "CONFIG_ERRORS <json>" with the
measured errors:
  plain-32x32x64-N50 : 50 1S iterations, FP64 and FP32 GPU vs the oracle;
  production-64x64x128-N100 : 2 iterations FP64 GPU vs the oracle, then
      200 iterations FP32 vs FP64 GPU;
  config3 sweep (48 systems, x = 56/64/72) : 40 iterations FP32 vs FP64 GPU,
      and the 72 x 64 x 128 system vs the oracle after 2 iterations;
  cl-96x96x192-N100 : 2 replicas, 150 Euler/CL steps FP32 vs FP64 GPU with
      the same noise, mean and variance of Re H over the last 100 steps.
"""
import json
import sys

import numpy as np
import pytest

import pyoracle
import scft_amd
from scft_amd import PREC_FP32, PREC_FP64, seeded_fields

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def report(name, d):
    say("CONFIG_ERRORS " + json.dumps({"config": name, **d}))


def rel(a, b):
    return float(abs(a - b) / abs(b))


def rel_arr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def gpu_state(ctx, sys_=0):
    rA, rB = ctx.get_density(sys=sys_)
    return ctx.calc_H()[sys_], ctx.get_Q(sys=sys_), rA, rB


def errors(got, want):
    H, Q, rA, rB = got
    Hr, Qr, rAr, rBr = want
    return {"dH_rel": rel(H, Hr), "dQ_rel": rel(Q, Qr), "drhoA_rel": rel_arr(rA, rAr),
            "drhoB_rel": rel_arr(rB, rBr)}


def oracle_state(o):
    return o.calc_H(), o.Q(), o.array("rhoda"), o.array("rhodb")


def test_errors_plain_32x32x64():
    w = scft_amd.config("plain-32x32x64-N50")
    p = w.params()
    f = seeded_fields(p, w.seed, period_rg=4.0, amp=1.0, sigma=0.01)
    n = 50
    say("plain-32x32x64-N50 oracle", n, "iterations")
    o = pyoracle.Oracle(p)
    o.set_fields(*f)
    o.calc_poly_density()
    for _ in range(n):
        o.update_1s()
    want = oracle_state(o)
    o.close()
    out = {"iterations": n}
    for prec, tag in ((PREC_FP64, "fp64"), (PREC_FP32, "fp32")):
        ctx = scft_amd.Context(p, precision=prec)
        ctx.set_fields(*f)
        ctx.calc_poly_density()
        ctx.iterate(n)
        out[tag] = errors(gpu_state(ctx), want)
        ctx.close()
    report("plain-32x32x64-N50", out)
    assert out["fp64"]["dH_rel"] < 1e-10 and out["fp64"]["dQ_rel"] < 1e-10
    assert out["fp64"]["drhoA_rel"] < 1e-9 and out["fp64"]["drhoB_rel"] < 1e-9
    # FP32 propagator: H within 1e-6; Q and the densities carry the
    # accumulated FP32 rounding of N steps (measured ~1e-5)
    assert out["fp32"]["dH_rel"] < 1e-6 and out["fp32"]["dQ_rel"] < 5e-5
    assert out["fp32"]["drhoA_rel"] < 5e-5 and out["fp32"]["drhoB_rel"] < 5e-5


def test_errors_production_64x64x128():
    w = scft_amd.config("production-64x64x128-N100")
    p = w.params()
    f = seeded_fields(p, w.seed, period_rg=4.0, amp=1.0, sigma=0.01)
    say("production-64x64x128-N100 oracle, 2 iterations")
    o = pyoracle.Oracle(p)
    o.set_fields(*f)
    o.calc_poly_density()
    for _ in range(2):
        o.update_1s()
    want = oracle_state(o)
    o.close()
    c64 = scft_amd.Context(p, precision=PREC_FP64)
    c32 = scft_amd.Context(p, precision=PREC_FP32)
    for c in (c64, c32):
        c.set_fields(*f)
        c.calc_poly_density()
        c.iterate(2)
    out = {"fp64_vs_oracle_2_iterations": errors(gpu_state(c64), want),
           "fp32_vs_oracle_2_iterations": errors(gpu_state(c32), want)}
    say("production-64x64x128-N100 FP32 vs FP64, 200 iterations")
    for c in (c64, c32):
        c.iterate(198)
    out["fp32_vs_fp64_200_iterations"] = errors(gpu_state(c32), gpu_state(c64))
    c64.close()
    c32.close()
    report("production-64x64x128-N100", out)
    e = out["fp64_vs_oracle_2_iterations"]
    assert e["dH_rel"] < 1e-10 and e["dQ_rel"] < 1e-10 and e["drhoA_rel"] < 1e-9
    assert out["fp32_vs_fp64_200_iterations"]["dH_rel"] < 1e-6


def test_errors_config3_sweep():
    ws = scft_amd.sweep("config3")
    ps = [w.params() for w in ws]
    fs = [seeded_fields(p, w.seed, period_rg=4.0, amp=1.0, sigma=0.01) for p, w in zip(ps, ws)]
    n = 40
    say("config3: 48 systems,", n, "iterations, FP32 and FP64")
    res = {}
    for prec in (PREC_FP64, PREC_FP32):
        ctx = scft_amd.Context(ps, precision=prec)
        for i, f in enumerate(fs):
            ctx.set_fields(*f, sys=i)
        ctx.calc_poly_density()
        ctx.iterate(2)
        k = 2  # the 72 x 64 x 128 system (x grid cycles 56 / 64 / 72)
        early = gpu_state(ctx, k)
        ctx.iterate(n - 2)
        H = ctx.calc_H()
        Q = np.array([ctx.get_Q(sys=i) for i in range(len(ps))])
        res[prec] = (early, H, Q)
        ctx.close()
    say("config3: oracle for the 72 x 64 x 128 system, 2 iterations")
    o = pyoracle.Oracle(ps[2])
    o.set_fields(*fs[2])
    o.calc_poly_density()
    for _ in range(2):
        o.update_1s()
    want = oracle_state(o)
    o.close()
    H64, H32 = res[PREC_FP64][1], res[PREC_FP32][1]
    Q64, Q32 = res[PREC_FP64][2], res[PREC_FP32][2]
    out = {"systems": len(ps), "iterations": n,
           "fp64_vs_oracle_72x64x128_2_iterations": errors(res[PREC_FP64][0], want),
           "fp32_vs_fp64_max_dH_rel": float(np.max(np.abs(H32 - H64) / np.abs(H64))),
           "fp32_vs_fp64_max_dQ_rel": float(np.max(np.abs(Q32 - Q64) / np.abs(Q64)))}
    report("config3", out)
    assert out["fp64_vs_oracle_72x64x128_2_iterations"]["dH_rel"] < 1e-10
    assert out["fp32_vs_fp64_max_dH_rel"] < 1e-6


def test_errors_cl_96x96x192():
    ws = [scft_amd.config("cl-96x96x192-N100", seed=2000 + r) for r in range(2)]
    ps = [w.params() for w in ws]
    f = seeded_fields(ps[0], 7, period_rg=4.0, amp=1.0, sigma=0.01)
    steps, keep = 150, 100
    say("cl-96x96x192-N100: 2 replicas,", steps, "steps, FP32 and FP64")
    series = {}
    for prec in (PREC_FP64, PREC_FP32):
        ctx = scft_amd.Context(ps, precision=prec)
        for r in range(2):
            ctx.set_fields(*f, sys=r)
        ctx.calc_poly_density()
        hs = []
        for s in range(steps):
            ctx.iterate(1)
            if s >= steps - keep:
                hs.append(ctx.calc_H())
        series[prec] = np.array(hs)
        ctx.close()
    h64, h32 = series[PREC_FP64].real, series[PREC_FP32].real
    out = {"replicas": 2, "steps": steps, "window": keep,
           "mean_H_fp64": h64.mean(axis=0).tolist(), "mean_H_fp32": h32.mean(axis=0).tolist(),
           "var_H_fp64": h64.var(axis=0).tolist(), "var_H_fp32": h32.var(axis=0).tolist(),
           "mean_dH_rel": float(np.max(np.abs(h32.mean(axis=0) - h64.mean(axis=0)) / np.abs(h64.mean(axis=0)))),
           "max_step_dH_rel": float(np.max(np.abs(h32 - h64) / np.abs(h64)))}
    report("cl-96x96x192-N100", out)
    assert out["mean_dH_rel"] < 1e-5
