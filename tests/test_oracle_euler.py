"""The oracle's Euler update (code/Euler_update.cpp:5-213), pinned by what
the mathematics fixes: without noise (do_CL = 0) one step moves each field
by -lambda times its full force (w+ at :58-67, the AB pair at :79-98), so
(w - w') / lambda must be dH/dw of the H the oracle evaluates
(code/calc_h.cpp), up to the cell volume: d/de H(w + e d) = dV <force, d>
along random directions d of every field.  This holds for chi_AB > 0.  The
negative-chi branch (sign flags of code/read_input.cpp:53-85) is not such a
gradient, as with the 1S update's negative-chi branches; its ratios are
reported and the mismatch asserted, so a change to that branch shows.  The
two step sizes differ, which pins which field steps with lam_imag and which
with lam_real.

Small 2D box with the published chemistry: first the w+ and AB fields alone,
then with chains grafted to the plates and the BC / AC pairs.
"""
import os

import numpy as np
import pytest

import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ["wpl", "wabp", "wabm"]


def small_ab(chi_ab, lam_imag=None, lam_real=None):
    p = pyoracle.read_input(os.path.join(GOLDEN, "bcp_dist-0.300.input"), 2)
    p.Nda, p.Ndb = 10, 10
    p.Nx[0], p.Nx[1] = 48, 80
    p.L[0], p.L[1] = 24.0, 40.0
    p.ic_flag[0] = 0
    p.ic_flag[1] = 0
    p.sigma = 0.0
    p.Ng = 0
    p.chi_bc = 0.0
    p.chi_ac = 0.0
    p.chi_ab = chi_ab
    p.update_scheme = 0
    p.do_CL = 0
    if lam_imag is not None:
        p.lam_imag, p.lam_real = lam_imag, lam_real
    return p


def _physical_unit(chi, k):
    """Direction of field k at a saddle point: w+ and wAB+ imaginary, wAB-
    real, the pair swapped for a negative chi (code/initialize.cpp:459-482)."""
    if k == 0:
        return 1j
    plus = k == 1
    if chi < 0:
        return 1.0 if plus else 1j
    return 1j if plus else 1.0


def _set(o, fields):
    for name, a in zip(FIELDS, fields):
        o.set_array(name, a)


@pytest.mark.parametrize("chi_ab", [0.501604, -0.3])
def test_euler_forces_are_derivatives_of_H(chi_ab):
    p = small_ab(chi_ab)
    rng = np.random.default_rng(9)
    f = [np.asarray(0.05 * rng.standard_normal(p.M) * _physical_unit(chi_ab, k), np.complex128)
         for k in range(3)]
    dirs = [rng.standard_normal(p.M) * _physical_unit(chi_ab, k) for k in range(3)]

    def H_at(fields):
        q = pyoracle.Oracle(p)
        _set(q, fields)
        q.calc_poly_density()
        h = q.calc_H()
        q.close()
        return h

    # w+ and wAB+ step with lam_imag, wAB- with lam_real
    # (code/Euler_update.cpp:64,84,95): distinct values pin the assignment
    lams = [1e-8, 1e-8, 3e-8]
    o = pyoracle.Oracle(small_ab(chi_ab, lam_imag=lams[0], lam_real=lams[2]))
    _set(o, f)
    o.calc_poly_density()
    o.update_euler()
    new = [o.array(n) for n in FIELDS]
    o.close()
    dV = (p.L[0] / p.Nx[0]) * (p.L[1] / p.Nx[1])
    for k in range(3):
        eps = 1e-6
        plus = [x.copy() for x in f]
        minus = [x.copy() for x in f]
        plus[k] = plus[k] + eps * dirs[k]
        minus[k] = minus[k] - eps * dirs[k]
        fd = (H_at(plus) - H_at(minus)) / (2 * eps)
        ratio = fd / np.sum((f[k] - new[k]) / lams[k] * dirs[k])
        print(FIELDS[k], chi_ab, ratio / dV)
        if k == 0 or chi_ab > 0:
            assert abs(ratio - dV) < 1e-5 * dV, (FIELDS[k], ratio, dV)
        else:
            assert abs(ratio - dV) > 1e-2 * dV   # the reference's non-gradient branch


ALL = ["wpl", "wabp", "wabm", "wbcp", "wbcm", "wacp", "wacm"]


def small_grafted(chi_bc, chi_ac, lam_imag=None, lam_real=None):
    p = small_ab(0.501604, lam_imag, lam_real)
    g = pyoracle.read_input(os.path.join(GOLDEN, "bcp_dist-0.300.input"), 2)
    p.sigma, p.Ng = g.sigma, 4
    p.chi_bc, p.chi_ac = chi_bc, chi_ac
    return p


def _unit7(p, k):
    """Saddle-point direction of field k (ALL order), swapped for chi < 0."""
    if k == 0:
        return 1j
    chi = [p.chi_ab, p.chi_ab, p.chi_bc, p.chi_bc, p.chi_ac, p.chi_ac][k - 1]
    plus = k % 2 == 1
    if chi < 0:
        return 1.0 if plus else 1j
    return 1j if plus else 1.0


@pytest.mark.parametrize("chi_bc", [0.30, -0.018209])
def test_euler_forces_with_grafts_and_pairs(chi_bc):
    # the same identity with chains grafted to the plates and the BC / AC
    # pairs (code/Euler_update.cpp:108-181): the graft density enters the w+
    # force (:61) and both pairs, and H carries -n ng Re Qga (calc_h.cpp:31-35)
    chi_ac = 0.25
    p = small_grafted(chi_bc, chi_ac)
    assert p.sigma > 0
    rng = np.random.default_rng(5)
    f = [np.asarray(0.05 * rng.standard_normal(p.M) * _unit7(p, k), np.complex128) for k in range(7)]
    dirs = [rng.standard_normal(p.M) * _unit7(p, k) for k in range(7)]

    def H_at(fields):
        q = pyoracle.Oracle(p)
        for name, a in zip(ALL, fields):
            q.set_array(name, a)
        q.calc_poly_density()
        h = q.calc_H()
        q.close()
        return h

    # + fields step with lam_imag, - fields with lam_real (:64,84,95,122,133,160,171)
    li, lr = 1e-8, 3e-8
    lams = [li, li, lr, li, lr, li, lr]
    o = pyoracle.Oracle(small_grafted(chi_bc, chi_ac, li, lr))
    for name, a in zip(ALL, f):
        o.set_array(name, a)
    o.calc_poly_density()
    o.update_euler()
    new = [o.array(n) for n in ALL]
    o.close()
    dV = (p.L[0] / p.Nx[0]) * (p.L[1] / p.Nx[1])
    chis = [None, p.chi_ab, p.chi_ab, p.chi_bc, p.chi_bc, p.chi_ac, p.chi_ac]
    for k in range(7):
        eps = 1e-6
        plus = [x.copy() for x in f]
        minus = [x.copy() for x in f]
        plus[k] = plus[k] + eps * dirs[k]
        minus[k] = minus[k] - eps * dirs[k]
        fd = (H_at(plus) - H_at(minus)) / (2 * eps)
        ratio = fd / np.sum((f[k] - new[k]) / lams[k] * dirs[k])
        print(ALL[k], chis[k], ratio / dV)
        if chis[k] is None or chis[k] > 0:
            assert abs(ratio - dV) < 1e-5 * dV, (ALL[k], ratio, dV)
        else:
            assert abs(ratio - dV) > 1e-3 * dV   # the reference's non-gradient branch
