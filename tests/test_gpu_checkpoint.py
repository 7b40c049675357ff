"""Checkpointed propagators (SCFT_OPT_CHECKPOINT): q kept at ceil(N/K)
checkpoints, K = ceil(sqrt N), and recomputed segment by segment while the
q-dagger chain advances.  The chains are the same kernels on the same inputs
as with full storage, so with the autotuner off (both modes then run the
same pass variants) Q and the densities agree to the FP64 summation order.
"""
import os
import re
import sys

import numpy as np
import pytest

import scft_amd
from scft_amd import PREC_FP32, PREC_FP64, Workload, seeded_fields

pytestmark = pytest.mark.gpu
CKPT = scft_amd.OPT_CHECKPOINT


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def contexts(ps, prec):
    os.environ["SCFT_NO_TUNE"] = "1"
    try:
        full = scft_amd.Context(ps, precision=prec)
        ck = scft_amd.Context(ps, precision=prec | CKPT)
    finally:
        del os.environ["SCFT_NO_TUNE"]
    return full, ck


@pytest.mark.parametrize("N,f_A,prec", [(16, 0.5, PREC_FP64), (50, 0.3, PREC_FP64), (50, 0.3, PREC_FP32),
                                        (7, 0.6, PREC_FP64)])
def test_checkpointed_density_equals_full_storage(N, f_A, prec):
    say("checkpoint", N, f_A, prec)
    ws = [Workload(grid=(16, 16, 32), box_rg=(2.0, 2.0, 4.0), N=N, f_A=f_A),
          Workload(grid=(24, 16, 32), box_rg=(3.0, 2.0, 4.0), N=N, f_A=f_A, d=0.25)]
    ps = [w.params() for w in ws]
    full, ck = contexts(ps, prec)
    for i, p in enumerate(ps):
        f = seeded_fields(p, 3 + i, period_rg=2.0)
        full.set_fields(*f, sys=i)
        ck.set_fields(*f, sys=i)
    full.calc_poly_density()
    ck.calc_poly_density()
    tol = 1e-13
    for i in range(2):
        assert abs(ck.get_Q(sys=i) - full.get_Q(sys=i)) <= tol * abs(full.get_Q(sys=i))
        a0, b0 = full.get_density(sys=i)
        a1, b1 = ck.get_density(sys=i)
        assert rel(a1, a0) < tol and rel(b1, b0) < tol, (rel(a1, a0), rel(b1, b0))
    # graph-captured iterations (the checkpointed stage forks and joins the
    # two streams once per segment)
    full.iterate(5)
    ck.iterate(5)
    H0, H1 = full.calc_H(), ck.calc_H()
    # FP64: equal to the summation order; FP32: the 1e-16 differences of the
    # densities feed the fields and can flip single FP32 roundings of exp(-w)
    tol_it = 1e-12 if prec == PREC_FP64 else 1e-7
    assert np.all(np.abs(H1 - H0) <= tol_it * np.abs(H0)), (H0, H1)
    # storage: 2 (K + 1) + ceil(N / K) slices instead of 2 N
    K = int(np.ceil(np.sqrt(N)))
    m = re.search(r"propagator_storage: (\d+) slices", ck.describe())
    assert int(m.group(1)) == 2 * (K + 1) + (N + K - 1) // K
    m = re.search(r"propagator_storage: (\d+) slices", full.describe())
    assert int(m.group(1)) == 2 * N
    full.close()
    ck.close()


def test_checkpointed_saddle_point_forces():
    # the forces need no stored q: at the converged saddle point they equal
    # those of the full-storage run
    say("checkpoint saddle forces")
    w = Workload(grid=(16, 16, 32), box_rg=(2.0, 2.0, 4.0), N=16, d=0.35, theta_deg=5.0, dt=0.5)
    p = w.params()
    f = seeded_fields(p, 4, period_rg=2.0, amp=1.0, sigma=0.0)
    out = []
    for prec in (PREC_FP64, PREC_FP64 | CKPT):
        ctx = scft_amd.Context(p, precision=prec)
        ctx.set_fields(*f)
        ctx.calc_poly_density()
        ctx.iterate(3000)
        out.append((ctx.calc_H()[0], ctx.plate_forces()))
        ctx.close()
    (H0, (D0, T0)), (H1, (D1, T1)) = out
    assert abs(H1 - H0) < 1e-11 * abs(H0)
    assert abs(D1 - D0) < 1e-8 * abs(D0) and abs(T1 - T0) < 1e-8 * abs(T0)


def test_checkpoint_rejects_unsupported():
    p = Workload(grid=(16, 16, 32), box_rg=(2.0, 2.0, 4.0), N=16, f_A=1.0).params()
    with pytest.raises(scft_amd.ScftError):
        scft_amd.Context(p, precision=PREC_FP64 | CKPT)
    a = Workload(grid=(16, 16, 32), box_rg=(2.0, 2.0, 4.0), N=16).params()
    b = Workload(grid=(16, 16, 32), box_rg=(2.0, 2.0, 4.0), N=20).params()
    with pytest.raises(scft_amd.ScftError):
        scft_amd.Context([a, b], precision=PREC_FP64 | CKPT)
