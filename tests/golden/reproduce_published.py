"""Runs published inputs as published, batched in one context, to the
reference's stop rule per run, and compares each run's last printed row with
its published LOG:

    python tests/golden/reproduce_published.py dist-0.140 dist-0.145 ... \\
        --out results.json [--max-seconds S] [--fp32]

Every run starts from its input's own init_fields flags (the published runs
restarted from .res fields that the snapshot does not hold).  Per run, as
simulate() does (code/simulate.cpp:85-157): a print every print_freq
iterations, error = |H - H_prev| / V / print_freq, close += (error <
10 error_tol), stop at the first print with iter > 25, error < error_tol and
close > 10.  Runs that stopped keep iterating in the batch; their printed
values at the stop are what is reported.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "aligned-nanoplate-separation-and-rotation_amd"))
sys.path.insert(0, HERE)

import scft_amd  # noqa: E402
from published_input import published  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("runs", nargs="+")
    ap.add_argument("--out", required=True)
    ap.add_argument("--max-seconds", type=float, default=1000.0)
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--forces", action="store_true",
                    help="also record dH/dc (plate centres, b units) and dH/dtheta (about z) at each stop")
    args = ap.parse_args()
    last = json.load(open(os.path.join(HERE, "log_last_rows.json")))
    ps, Hpub = [], []
    with tempfile.TemporaryDirectory() as td:
        for run in args.runs:
            text, h = published(run)
            path = os.path.join(td, "bcp.input")
            open(path, "w").write(text)
            ps.append(scft_amd.read_input(path, 2))
            Hpub.append(h)
    pf = ps[0].print_freq
    assert all(p.print_freq == pf for p in ps), "one print_freq per batch"
    t0 = time.time()
    ctx = scft_amd.Context(ps, precision=scft_amd.PREC_FP32 if args.fp32 else scft_amd.PREC_FP64)
    print(f"{len(ps)} runs, context ready in {time.time() - t0:.1f} s", flush=True)
    infos = [ctx.info(i) for i in range(len(ps))]
    ctx.calc_poly_density()
    n = len(ps)
    Hloc = np.zeros(n, complex)
    close = np.zeros(n, int)
    stop = [None] * n
    it = 0
    ctx.iterate(1)   # iteration 0
    while True:
        H = ctx.calc_H()
        err = np.abs(H - Hloc) / np.array([f.V for f in infos]) / pf
        Hloc = H
        for i in range(n):
            if stop[i] is not None:
                continue
            if err[i] < ps[i].error_tol * 10:
                close[i] += 1
            if it > 25 and err[i] < ps[i].error_tol and close[i] > 10:
                Q = ctx.get_Q(sys=i)
                Qg = ctx.get_Qga(sys=i) if ps[i].sigma > 0 else 0.0
                stop[i] = dict(run=args.runs[i], iter=it, H=H[i].real, mlogQd=-math.log(abs(Q)),
                               mlogQg=-ps[i].n_exp_nr * infos[i].ng_per_np * Qg.real, err=float(err[i]),
                               H_published=Hpub[i], published_last_row=last[args.runs[i]][-1])
                if args.forces:
                    g = ctx.grad_plates(axis=(0.0, 0.0, 1.0), sys=i)
                    stop[i]["dH_dc"] = [[g[j, k].real for k in range(3)] for j in range(ps[i].n_exp_nr)]
                    stop[i]["dH_dtheta"] = [g[j, 3].real for j in range(ps[i].n_exp_nr)]
                    stop[i]["centres"] = [[ps[i].exp_nr_c[j][k] * ps[i].L[k] for k in range(2)]
                                          for j in range(ps[i].n_exp_nr)]
                print(f" {args.runs[i]} stopped at {it}: H {H[i].real:.6f} published {Hpub[i]}", flush=True)
        if all(s is not None for s in stop) or time.time() - t0 > args.max_seconds:
            break
        ctx.iterate(pf)
        it += pf
        if (it // pf) % 20 == 0:
            print(f" iteration {it}, {sum(s is not None for s in stop)} of {n} stopped, "
                  f"{time.time() - t0:.0f} s", flush=True)
    res = dict(iterations=it, seconds=time.time() - t0, precision="fp32" if args.fp32 else "fp64",
               runs=[s if s is not None else dict(run=args.runs[i], iter=None, H=Hloc[i].real,
                                                  H_published=Hpub[i], err=float(err[i]))
                     for i, s in enumerate(stop)])
    json.dump(res, open(args.out, "w"), indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
