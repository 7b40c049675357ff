"""CPU checks of the workload builders (units, plates, sweeps)."""
import math

import numpy as np
import pytest

import scft_amd


def test_units_mapping():
    w = scft_amd.config("production-64x64x128-N100")
    p = w.params()
    rg = math.sqrt(99 / 6)
    assert p.dim == 3 and (p.Nx[0], p.Nx[1], p.Nx[2]) == (64, 64, 128)
    assert p.L[0] == pytest.approx(4 * rg) and p.L[2] == pytest.approx(8 * rg)
    assert p.chi_ab == pytest.approx(0.2) and p.kappa == pytest.approx(0.5)
    assert p.rho0 == pytest.approx(50 * 100 / rg ** 3)
    assert p.lam_imag == pytest.approx(0.005 / (p.rho0 * 100))
    assert (p.Nda, p.Ndb) == (50, 50)
    # plates: thickness 0.2 Rg, half side 0.5 Rg, interface 0.05 Rg, d = 0.30
    assert p.L_nr == pytest.approx(0.2 * rg) and p.R_nr == pytest.approx(0.5 * rg)
    assert p.exp_nr_c[0][0] == pytest.approx(0.35) and p.exp_nr_c[1][0] == pytest.approx(0.65)
    assert p.exp_nr_u[0][2] == pytest.approx(1.0)


def test_config3_sweep():
    ws = scft_amd.sweep("config3")
    assert len(ws) == 48
    ds = [w.d for w in ws[:33]]
    assert ds[0] == pytest.approx(0.14) and ds[-1] == pytest.approx(0.46)
    assert np.allclose(np.diff(ds), 0.01)
    assert all(w.theta_deg == 0.0 for w in ws[:33])
    angles = [w.theta_deg for w in ws[33:]]
    assert angles == [float(a) for a in list(range(11)) + [20, 30, 45, 90]]
    assert all(w.d == 0.30 for w in ws[33:])
    assert [w.grid[0] for w in ws[:6]] == [56, 64, 72, 56, 64, 72]
    assert [w.box_rg[0] for w in ws[:3]] == [3.5, 4.0, 4.5]
    # rotation about y: u = (sin t, 0, cos t)
    p = ws[33 + 13].params()   # 45 deg
    assert p.exp_nr_u[0][0] == pytest.approx(math.sqrt(0.5))
    assert p.exp_nr_u[0][2] == pytest.approx(math.sqrt(0.5))


def test_headline_batch():
    ws = scft_amd.sweep("separation-64")
    assert len(ws) == 16
    assert all(w.grid == (64, 64, 128) and w.N == 100 for w in ws)
    assert len({w.d for w in ws}) == 16
