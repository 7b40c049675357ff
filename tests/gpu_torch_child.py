"""Child process of tests/test_gpu_torch.py: the libtorch plate-energy
wrapper on the GPU.  torch is imported first so that the process holds one
HIP runtime (torch's), which libscft_amd.so then shares.  Prints one line
"TORCH_RESULTS <json>" and exits non-zero on a failed check."""
import json
import os
import sys

import torch  # first

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "aligned-nanoplate-separation-and-rotation_amd"))
import numpy as np  # noqa: E402

import scft_amd  # noqa: E402
import scft_torch  # noqa: E402
from scft_amd import Workload, seeded_fields  # noqa: E402


def main():
    w = Workload(grid=(16, 16, 32), box_rg=(2.0, 2.0, 4.0), N=16, d=0.40, theta_deg=10.0, dt=0.5)
    p = w.params()
    f = seeded_fields(p, 4, period_rg=2.0, amp=1.0, sigma=0.0)
    ctx = scft_amd.Context(p)
    ctx.set_fields(*f)
    ctx.calc_poly_density()
    Lx = p.L[0]
    D0, th0 = w.d * Lx, np.radians(w.theta_deg)
    out = {}
    # fixed fields: exact derivative of H at these fields
    geom = torch.tensor([D0, th0], dtype=torch.float64, requires_grad=True)
    H = scft_torch.plate_energy(ctx, geom)
    H.backward()
    g = geom.grad.clone()

    def E(D, th, **kw):
        with torch.no_grad():
            return scft_torch.plate_energy(ctx, torch.tensor([D, th], dtype=torch.float64), **kw).item()

    hD, hT = 1e-3 * w.rg, np.radians(0.1)
    rich = []
    for k, h in ((0, hD), (1, hT)):
        e = [np.zeros(2), np.zeros(2)]
        for j, s in enumerate((1.0, 2.0)):
            v = np.array([D0, th0])
            dv = np.zeros(2)
            dv[k] = s * h
            e[j] = (E(*(v + dv)) - E(*(v - dv))) / (2 * s * h)
        rich.append((4 * e[0] - e[1]) / 3)
    out["fixed_fields"] = {"H": H.item(), "grad": g.tolist(), "richardson_fd": [float(r) for r in rich]}
    ok = all(abs(g[k].item() - rich[k]) < 1e-6 * abs(rich[k]) for k in range(2))
    # torch.autograd.gradcheck on the fixed-field energy
    gc = torch.autograd.gradcheck(lambda x: scft_torch.plate_energy(ctx, x), (geom.detach().clone().requires_grad_(True),),
                                  eps=1e-5, atol=1e-4, rtol=1e-5, raise_exception=False)
    out["gradcheck_fixed_fields"] = bool(gc)
    ok = ok and gc
    # saddle point: relax at each geometry; the gradient is the force/torque
    geom = torch.tensor([D0, th0], dtype=torch.float64, requires_grad=True)
    H = scft_torch.plate_energy(ctx, geom, max_iter=20000, tol=1e-13)
    H.backward()
    g = geom.grad.clone()
    fds = []
    for k, h in ((0, hD), (1, hT)):
        e = []
        for s in (1.0, 2.0):
            v = np.array([D0, th0])
            dv = np.zeros(2)
            dv[k] = s * h
            e.append((E(*(v + dv), max_iter=20000, tol=1e-13) - E(*(v - dv), max_iter=20000, tol=1e-13)) / (2 * s * h))
        fds.append((4 * e[0] - e[1]) / 3)
    out["saddle"] = {"H": H.item(), "grad": g.tolist(), "richardson_fd": [float(x) for x in fds]}
    ok = ok and all(abs(g[k].item() - fds[k]) < 1e-4 * abs(fds[k]) for k in range(2))
    out["hip_runtimes"] = sorted(set(l.split()[-1] for l in open("/proc/self/maps").read().splitlines()
                                     if "libamdhip64" in l))
    ctx.close()
    print("TORCH_RESULTS " + json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
