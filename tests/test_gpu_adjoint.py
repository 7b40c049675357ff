"""Plate forces and torques at converged saddle points of the 32x32x64, N = 50
system: d in {0.20, 0.30, 0.40} x theta in {0, 5, 45} deg, with the
propagators checkpointed every ceil(sqrt N) = 8 contour steps.

For each geometry the batch holds the centre system and the systems displaced
by +-delta and +-2 delta in D (delta = 1e-3 Rg) and in theta (0.1 deg), 81
systems in one FP64 context, all converged from the same lamellar start.
Checks:
  * the analytic dH/dD and dH/dtheta at the centre's saddle point
    (scft_grad_plates) against central differences of the converged H*
    (Richardson extrapolation of the delta and 2 delta differences), to
    1e-4 relative;
  * the same analytic derivative against the oracle's: central differences of
    the oracle's H at the centre's converged fields with displaced plates.
"""
import json
import sys

import numpy as np
import pytest

import pyoracle
import scft_amd
from scft_amd import Workload, seeded_fields

pytestmark = pytest.mark.gpu

GEOMS = [(d, th) for d in (0.20, 0.30, 0.40) for th in (0.0, 5.0, 45.0)]
SHIFTS = [("c", 0), ("d", -2), ("d", -1), ("d", 1), ("d", 2), ("t", -2), ("t", -1), ("t", 1), ("t", 2)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def base_workload(d, th):
    return Workload(grid=(32, 32, 64), box_rg=(4.0, 4.0, 8.0), N=50, d=d, theta_deg=th, dt=0.5)


def test_saddle_point_forces_and_torques():
    w0 = base_workload(0.3, 0.0)
    Lx = w0.params().L[0]
    dD = 1e-3 * w0.rg          # b units
    dth = 0.1                  # degrees
    ws = []
    for d, th in GEOMS:
        for kind, k in SHIFTS:
            dd = d + (k * dD / Lx if kind == "d" else 0.0)
            tt = th + (k * dth if kind == "t" else 0.0)
            ws.append(base_workload(dd, tt))
    ps = [w.params() for w in ws]
    f = seeded_fields(ps[0], 11, period_rg=4.0, amp=1.0, sigma=0.0)
    say("adjoint: 9 geometries x 9 systems, FP64, checkpointed propagators")
    # q checkpointed every K = ceil(sqrt N) contour steps (SCFT_OPT_CHECKPOINT)
    ctx = scft_amd.Context(ps, precision=scft_amd.PREC_FP64 | scft_amd.OPT_CHECKPOINT)
    say(" " + [l for l in ctx.describe().splitlines() if l.startswith("propagator_storage")][0])
    for i in range(len(ps)):
        ctx.set_fields(*f, sys=i)
    ctx.calc_poly_density()
    prev = ctx.calc_H()
    done = 0
    for chunk in range(200):
        ctx.iterate(100)
        done += 100
        H = ctx.calc_H()
        change = float(np.max(np.abs(H - prev) / np.abs(H)))
        prev = H
        if chunk % 10 == 0:
            say(" iterations", done, "max relative change of H over 100:", change)
        if change < 2e-14:
            break
    say(" converged after", done, "iterations, change", change)
    assert change < 1e-12
    out = []
    nS = len(SHIFTS)
    for g, (d, th) in enumerate(GEOMS):
        h = {s: H[g * nS + i].real for i, s in enumerate(SHIFTS)}
        fd1 = (h[("d", 1)] - h[("d", -1)]) / (2 * dD)
        fd2 = (h[("d", 2)] - h[("d", -2)]) / (4 * dD)
        rD = (4 * fd1 - fd2) / 3
        r = np.radians(dth)
        ft1 = (h[("t", 1)] - h[("t", -1)]) / (2 * r)
        ft2 = (h[("t", 2)] - h[("t", -2)]) / (4 * r)
        rT = (4 * ft1 - ft2) / 3
        aD, aT = ctx.plate_forces(sys=g * nS)
        out.append(dict(d=d, theta=th, dH_dD=aD.real, fd_dD=fd1, richardson_dD=rD, dH_dtheta=aT.real,
                        fd_dtheta=ft1, richardson_dtheta=rT))
        say(f" d={d:.2f} th={th:4.1f}: dH/dD {aD.real:.8e} fd {fd1:.8e} R {rD:.8e} | "
            f"dH/dth {aT.real:.8e} fd {ft1:.8e} R {rT:.8e}")
    # oracle gradient at the converged centre fields of three geometries
    for g in (0, 4, 8):
        d, th = GEOMS[g]
        fields = ctx.get_fields(sys=g * nS)

        def Ho(dd, tt):
            o = pyoracle.Oracle(base_workload(dd, tt).params())
            o.set_fields(*fields)
            o.calc_poly_density()
            v = o.calc_H().real
            o.close()
            return v
        o1 = (Ho(d + dD / Lx, th) - Ho(d - dD / Lx, th)) / (2 * dD)
        o2 = (Ho(d + 2 * dD / Lx, th) - Ho(d - 2 * dD / Lx, th)) / (4 * dD)
        out[g]["oracle_richardson_dD_fixed_fields"] = (4 * o1 - o2) / 3
        t1 = (Ho(d, th + dth) - Ho(d, th - dth)) / (2 * np.radians(dth))
        t2 = (Ho(d, th + 2 * dth) - Ho(d, th - 2 * dth)) / (4 * np.radians(dth))
        out[g]["oracle_richardson_dtheta_fixed_fields"] = (4 * t1 - t2) / 3
    ctx.close()
    say("ADJOINT_RESULTS " + json.dumps(out))
    scaleD = max(abs(o["richardson_dD"]) for o in out)
    scaleT = max(abs(o["richardson_dtheta"]) for o in out)
    for o in out:
        # relative to the derivative itself, or to the largest one of the
        # sweep where it vanishes by symmetry (theta = 0)
        assert abs(o["dH_dD"] - o["richardson_dD"]) < 1e-4 * max(abs(o["richardson_dD"]), 1e-3 * scaleD), o
        assert abs(o["dH_dtheta"] - o["richardson_dtheta"]) < 1e-4 * max(abs(o["richardson_dtheta"]),
                                                                       1e-3 * scaleT), o
        if "oracle_richardson_dD_fixed_fields" in o:
            assert abs(o["dH_dD"] - o["oracle_richardson_dD_fixed_fields"]) < 1e-6 * scaleD, o
            assert abs(o["dH_dtheta"] - o["oracle_richardson_dtheta_fixed_fields"]) < 1e-6 * scaleT, o
