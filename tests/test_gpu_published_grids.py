"""The published 2D grids (525 x 875 separation runs, 625 x 675 angle runs;
radices 3, 5, 7) through the odd-length line passes, against the oracle.

The published inputs (tests/golden/bcp_dist-0.300.input,
bcp_angle-45.00.input) have grafted chains, which this path does not build;
the tests run them with the grafts removed (sigma = 0, chi_BC = chi_AC = 0)
and everything else as published: N = 60, a = 0.2, rho0 = 10.0566,
chi_AB = 0.501604, kappa = 50, the plates and the 1S steps.
"""
import os
import sys
import time

import numpy as np
import pytest

import pyoracle
import scft_amd

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def graft_free(name):
    p = scft_amd.read_input(os.path.join(GOLDEN, name), 2)
    p.sigma = 0.0
    p.Ng = 0
    p.chi_bc = 0.0
    p.chi_ac = 0.0
    p.seed = 1
    return p


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("name", ["bcp_dist-0.300.input", "bcp_angle-45.00.input"])
def test_published_grid_iteration_matches_oracle(name):
    p = graft_free(name)
    say(name, "grid", p.Nx[0], p.Nx[1], "N", p.N)
    ctx = scft_amd.Context(p)
    ref = pyoracle.Oracle(p)
    info = ctx.info()
    for k in ("V", "V_1_exp_np", "np_frac", "phiD", "nD"):
        assert abs(getattr(info, k) - ref.scalar(k)) <= 1e-12 * abs(ref.scalar(k)), k
    # FFT seam at the published size
    rng = np.random.default_rng(2)
    x = rng.standard_normal(p.M) + 1j * rng.standard_normal(p.M)
    assert rel(ctx.fft(x, -1), ref.fft(x, -1)) < 1e-13
    assert rel(ctx.fft(x, +1), ref.fft(x, +1)) < 1e-13
    # the start fields of the input's own flags (lamellae masked by the plates)
    f = ctx.get_fields()
    fo = [ref.array(n) for n in ("wpl", "wabp", "wabm")]
    for a, b in zip(f, fo):
        assert np.max(np.abs(a - b)) <= 1e-14 * max(np.max(np.abs(b)), 1.0)
    ctx.calc_poly_density()
    ref.calc_poly_density()
    assert abs(ctx.get_Q() - ref.Q()) <= 1e-11 * abs(ref.Q())
    for it in range(2):
        ctx.update_1s()
        ref.update_1s()
    H, Hr = ctx.calc_H()[0], ref.calc_H()
    say(" H after 2 iterations", H, Hr, abs(H - Hr) / abs(Hr))
    assert abs(H - Hr) <= 1e-10 * abs(Hr)
    rA, _ = ctx.get_density()
    assert rel(rA, ref.array("rhoda")) < 1e-10
    # rate of the graph-captured loop at the published size
    ctx.iterate(2)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.iterate(20)
    ctx.synchronize()
    say(f" {name}: {20 / (time.perf_counter() - t0):.1f} iter/s (FP64)")
    ctx.close()
    ref.close()


@pytest.mark.parametrize("name", ["bcp_dist-0.300.input", "bcp_angle-45.00.input"])
@pytest.mark.parametrize("pick", [0, 1, 2, 3, 4, 5])
def test_every_odd_line_variant(name, pick):
    # SCFT_TUNE_PICK fixes the variant of every multi-variant launch (the
    # last one for picks beyond a launch's count): each two-step and
    # Stockham shape of the x and y passes, in FP64 and FP32, through the FFT
    # seam and one density evaluation
    p = graft_free(name)
    say(name, "variant pick", pick)
    ref = pyoracle.Oracle(p)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(p.M) + 1j * rng.standard_normal(p.M)
    ref.calc_poly_density()
    os.environ["SCFT_TUNE_PICK"] = str(pick)
    try:
        for prec, tol in ((scft_amd.PREC_FP64, 1e-12), (scft_amd.PREC_FP32, 2e-5)):
            ctx = scft_amd.Context(p, precision=prec)
            plan = ctx.describe()
            assert "odd_line[" in plan, plan
            if prec == scft_amd.PREC_FP64:
                assert rel(ctx.fft(x, -1), ref.fft(x, -1)) < 1e-13
                assert rel(ctx.fft(x, +1), ref.fft(x, +1)) < 1e-13
            ctx.calc_poly_density()
            dq = abs(ctx.get_Q() - ref.Q()) / abs(ref.Q())
            rA, rB = ctx.get_density()
            da, db = rel(rA, ref.array("rhoda")), rel(rB, ref.array("rhodb"))
            say(f" prec {prec}: dQ/Q {dq:.2e} drhoA {da:.2e} drhoB {db:.2e}")
            assert dq < tol and da < tol and db < tol
            ctx.close()
    finally:
        del os.environ["SCFT_TUNE_PICK"]
    ref.close()
