"""Pins the FP64 CPU oracle before the library is compared against it.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed operand) fails it:
  * FFT: numpy.fft golden vectors (pocketfft, not the reference), a naive DFT,
    Parseval, and the forward 1/M scale of fft_fwd_wrapper
    (code/fft_wrappers.cpp:23,28);
  * plate density + volume bookkeeping: every init-time scalar printed in the
    52 published LOGs (V_1_exp_np, nD, np_frac, phiG, Vseg, nD*N);
  * propagator / Q / density / 1S / H: closed forms and exact invariants.
"""
import json
import math
import os
import tempfile

import numpy as np
import pytest

import pyoracle as po
from scft_amd import Workload, seeded_fields


@pytest.fixture(scope="module", autouse=True)
def _threads():
    po.set_threads(min(8, os.cpu_count() or 1))


# ---------------------------------------------------------------- FFT
def test_fft1d_matches_numpy_golden(golden_dir):
    g = np.load(os.path.join(golden_dir, "fft_vectors.npz"))
    for key in g.files:
        if not key.startswith("x"):
            continue
        n = key[1:]
        x = g[key]
        F = po.fft1d(x, -1)
        B = po.fft1d(x, +1)
        scale = np.max(np.abs(g["F" + n]))
        assert np.max(np.abs(F - g["F" + n])) <= 1e-14 * scale, n
        assert np.max(np.abs(B - g["B" + n])) <= 1e-14 * scale, n


def test_fft3d_matches_numpy_golden(golden_dir):
    g = np.load(os.path.join(golden_dir, "fft_vectors.npz"))
    for key in g.files:
        if not key.startswith("a"):
            continue
        a = g[key]
        A = po.fft3d(a, -1)
        assert np.max(np.abs(A - g["A" + key[1:]])) <= 1e-13 * np.max(np.abs(A))


def test_fft_naive_dft_and_parseval():
    rng = np.random.default_rng(1)
    for n in (6, 10, 21, 35, 49):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        j = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(j, j) / n)
        F = po.fft1d(x, -1)
        assert np.allclose(F, W @ x, atol=1e-12)
        assert math.isclose(np.sum(abs(F) ** 2), n * np.sum(abs(x) ** 2), rel_tol=1e-13)
        assert np.allclose(po.fft1d(F, +1) / n, x, atol=1e-14)


def test_fft_wrapper_scaling_and_plane_wave():
    w = Workload(grid=(8, 6, 10), box_rg=(2.0, 2.0, 3.0), N=8, n_plates=0)
    o = po.Oracle(w.params())
    nx, ny, nz = 8, 6, 10
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky, kz = 3, 1, 7
    a = np.exp(2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz)).reshape(-1)
    F = o.fft(a, -1)           # forward, scaled 1/M: a delta of height 1
    idx = kx + nx * (ky + ny * kz)
    expect = np.zeros_like(F)
    expect[idx] = 1.0
    assert np.max(np.abs(F - expect)) < 1e-14
    assert np.max(np.abs(o.fft(F, +1) - a)) < 1e-13   # backward unscaled


# ------------------------------------------------ known answers of the LOGs
def _runs(golden_dir):
    return json.load(open(os.path.join(golden_dir, "log_known_answers.json")))["runs"]


def _check_printed(value, printed):
    # the LOG prints with %lf: six decimals, rounded
    assert abs(value - printed) <= 5.01e-7 + 1e-12 * abs(printed), (value, printed)


def test_log_known_answers(golden_dir):
    runs = _runs(golden_dir)
    assert len(runs) == 52
    with tempfile.TemporaryDirectory() as td:
        for r in runs:
            path = os.path.join(td, "bcp.input")
            open(path, "w").write(r["input"])
            p = po.read_input(path, 2)
            o = po.Oracle(p)
            log = r["log"]
            _check_printed(o.scalar("V_1_exp_np"), log["V_1_exp_np"])
            _check_printed(o.scalar("nD"), log["nD"])
            _check_printed(o.scalar("np_frac"), log["np_frac"])
            _check_printed(o.scalar("phiG"), log["phiG"])
            _check_printed(o.scalar("ng_per_np"), log["ng_per_np"])
            _check_printed(o.scalar("V"), log["V"])
            N = o.scalar("N")
            vseg = o.scalar("nD") * N + o.scalar("nP") * o.scalar("ng_per_np") * p.Ng
            _check_printed(vseg, log["Vseg"])
            _check_printed(o.scalar("nD") * N, log["nDN"])
            o.close()


def test_parser_reads_published_input(golden_dir):
    p = po.read_input(os.path.join(golden_dir, "bcp_dist-0.300.input"), 2)
    assert (p.Nda, p.Ndb) == (30, 30)
    assert (p.Nx[0], p.Nx[1]) == (525, 875)
    assert (p.L[0], p.L[1]) == (52.0, 90.0)
    assert p.chi_bc == pytest.approx(-0.018209)
    assert p.lam_imag == 0.005 and p.lam_real == 0.001
    assert p.n_exp_nr == 2 and p.np_type == 2 and p.Ng == 9
    assert p.exp_nr_c[0][1] == pytest.approx(0.65) and p.exp_nr_c[1][1] == pytest.approx(0.35)
    q = po.read_input(os.path.join(golden_dir, "bcp_angle-45.00.input"), 2)
    assert q.exp_nr_u[0][0] == pytest.approx(math.sqrt(0.5))
    assert q.exp_nr_u[0][0] ** 2 + q.exp_nr_u[0][1] ** 2 == pytest.approx(1.0, abs=1e-15)
    # a 3D read takes the third number of each grid / box / centre line
    r = po.read_input(os.path.join(golden_dir, "bcp_dist-0.300.input"), 3)
    assert (r.Nx[0], r.Nx[1], r.Nx[2]) == (525, 875, 875)
    assert r.L[2] == 25.0


# -------------------------------------------------- kernels: closed forms
def _small(**kw):
    base = dict(grid=(8, 8, 16), box_rg=(2.0, 2.0, 4.0), N=8, n_plates=0)
    base.update(kw)
    return Workload(**base)


def _kgrid(p):
    nx, ny, nz = p.grid()
    Lx, Ly, Lz = p.box()
    k = []
    for n, L in ((nx, Lx), (ny, Ly), (nz, Lz)):
        m = np.arange(n)
        k.append(2 * np.pi * np.where(m < n / 2, m, m - n) / L)
    KZ, KY, KX = np.meshgrid(k[2], k[1], k[0], indexing="ij")
    return (KX ** 2 + KY ** 2 + KZ ** 2).reshape(-1)


def test_kernels_closed_forms():
    w = _small()
    p = w.params()
    o = po.Oracle(p)
    k2 = _kgrid(p)
    assert np.allclose(o.array("bond").real, np.exp(-k2 / 6.0), rtol=1e-15, atol=0)
    assert np.allclose(o.array("hhat").real, np.exp(-k2 * p.a_smear ** 2 / 2.0), rtol=1e-15)
    f = p.Nda / p.N
    x = k2 * (p.N - 1) / 6.0
    with np.errstate(divide="ignore", invalid="ignore"):
        gaa = 2 * (np.exp(-x * f) + f * x - 1) / x ** 2
        gab = (1 - np.exp(-x * f)) * (1 - np.exp(-x * (1 - f))) / x ** 2
    gaa[0], gab[0] = f * f, f * (1 - f)
    assert np.allclose(o.array("gaa").real, gaa, rtol=1e-13)
    assert np.allclose(o.array("gab").real, gab, rtol=1e-13)
    # k=0 limits are the x->0 limits of the continuous formulas
    xs = 1e-4
    assert o.array("gaa")[0].real == pytest.approx(2 * (math.exp(-xs * f) + f * xs - 1) / xs ** 2, rel=1e-3)


def test_square_plate_volume_continuum():
    # a well-resolved square plate: integral -> thickness * side^2
    w = Workload(grid=(64, 64, 64), box_rg=(2.0, 2.0, 2.0), N=8, n_plates=1,
                 plate_side_rg=1.0, plate_thick_rg=0.5, plate_xi_rg=0.02, d=0.0)
    p = w.params()
    o = po.Oracle(p)
    rg = w.rg
    assert o.scalar("V_1_exp_np") == pytest.approx(0.5 * 1.0 * rg ** 3, rel=2e-3)


# ------------------------------------------- propagator / Q / density
def test_propagator_zero_and_constant_fields():
    w = _small()
    p = w.params()
    o = po.Oracle(p)
    M = p.M
    z = np.zeros(M, complex)
    o.set_fields(z, z, z)
    o.calc_poly_density()
    assert abs(o.Q() - 1.0) < 1e-14
    for j in (0, 3, p.N - 1):
        assert np.max(np.abs(o.array(f"q:{j}") - 1.0)) < 1e-13
    # constant w+ = i c  ->  wA = wB = -c  ->  Q = exp(N c)
    c = 0.37
    o.set_fields(np.full(M, 1j * c), z, z)
    o.calc_poly_density()
    assert abs(o.Q() - math.exp(p.N * c)) < 1e-12 * math.exp(p.N * c)
    # uniform densities: rhoA = nD Na / V, rhoB = nD Nb / V
    rA, rB = o.array("rhoda"), o.array("rhodb")
    nD, V = o.scalar("nD"), o.scalar("V")
    assert np.allclose(rA, nD * p.Nda / V, rtol=1e-13)
    assert np.allclose(rB, nD * p.Ndb / V, rtol=1e-13)


def test_propagator_plane_wave_decay():
    # w = 0 except q0 = 1 + eps cos(k x): each step multiplies the mode by
    # exp(-k^2/6); checked through one explicit q step of the oracle.
    w = _small(N=4)
    p = w.params()
    o = po.Oracle(p)
    M = p.M
    z = np.zeros(M, complex)
    nx, ny, nz = p.grid()
    Lx = p.box()[0]
    x = (np.arange(M) % nx) * (Lx / nx)
    k = 2 * np.pi * 2 / Lx
    eps = 1e-3
    # w+ = i log(1 + eps cos kx) makes wA = i w+ = -log(...), so q0 = 1 + eps cos
    wpl = 1j * np.log(1 + eps * np.cos(k * x))
    # smearing would blur it: use a = 0 for this check
    p.a_smear = 0.0
    o = po.Oracle(p)
    o.set_fields(wpl, z, z)
    o.calc_poly_density()
    q0 = o.array("q:0")
    q1 = o.array("q:1")
    assert np.allclose(q0, 1 + eps * np.cos(k * x), atol=1e-14)
    expect1 = (1 + eps * np.exp(-k * k / 6) * np.cos(k * x)) * (1 + eps * np.cos(k * x))
    assert np.allclose(q1, expect1, atol=1e-13)


def test_density_sum_rule_random_fields():
    w = _small(n_plates=2, np_type=3)
    p = w.params()
    o = po.Oracle(p)
    wpl, wabp, wabm = seeded_fields(p, 7, period_rg=2.0, amp=0.5, sigma=0.1)
    o.set_fields(wpl, wabp, wabm)
    o.calc_poly_density()
    s = o.integ(o.array("rhoda") + o.array("rhodb"))
    nDN = o.scalar("nD") * p.N
    assert abs(s - nDN) < 1e-11 * nDN


def test_swap_symmetry_qdag():
    # f_A = 1/2: swapping A and B fields swaps q and q-dagger
    w = _small(N=8)
    p = w.params()
    o = po.Oracle(p)
    wpl, wabp, wabm = seeded_fields(p, 3, period_rg=2.0, amp=0.5, sigma=0.1)
    o.set_fields(wpl, wabp, wabm)
    o.calc_poly_density()
    q5 = o.array("q:5")
    qd5 = o.array("qdag:5")
    o.set_fields(wpl, wabp, -wabm)   # wAB- -> -wAB- swaps wA and wB
    o.calc_poly_density()
    assert np.allclose(o.array("qdag:5"), q5, rtol=1e-12, atol=1e-14)
    assert np.allclose(o.array("q:5"), qd5, rtol=1e-12, atol=1e-14)


# --------------------------------------------------------------- H, 1S
def test_H_uniform_closed_form():
    w = _small()
    p = w.params()
    o = po.Oracle(p)
    M = p.M
    z = np.zeros(M, complex)
    c = 0.21
    o.set_fields(np.full(M, 1j * c), z, z)
    o.calc_poly_density()
    H = o.calc_H()
    V = o.scalar("V")
    expect = -V * p.rho0 * c * c / (2 * p.kappa)   # the i rho0 w+ and -nD log Q terms cancel
    assert abs(H - expect) < 1e-10 * abs(expect)


def test_1s_uniform_mode_relaxation():
    # k = 0 of the w+ update: F = 0 at the uniform solution, A - B = rho0/kappa,
    # so the mean relaxes by (1 + lam B0)/(1 + lam A0), A0 = rho0/kappa + rho0 N.
    w = _small()
    p = w.params()
    o = po.Oracle(p)
    M = p.M
    z = np.zeros(M, complex)
    c = 0.05
    o.set_fields(np.full(M, 1j * c), z, z)
    o.calc_poly_density()
    o.update_1s()
    wpl, wabp, wabm = o.array("wpl"), o.array("wabp"), o.array("wabm")
    lam = p.lam_imag
    A0 = p.rho0 / p.kappa + p.rho0 * p.N
    B0 = p.rho0 * p.N
    assert np.allclose(wpl, 1j * c * (1 + lam * B0) / (1 + lam * A0), rtol=1e-12)
    # wAB+ has no k=0 delta in its force (code/1s_update.cpp:101): F = i rho0,
    # A = 2 rho0/chi + rho0 N, B = rho0 N, from wAB+ = 0
    Aab = 2 * p.rho0 / p.chi_ab + p.rho0 * p.N
    assert np.allclose(wabp, -1j * lam * p.rho0 / (1 + lam * Aab), rtol=1e-12)
    # wAB- force h (rhoB - rhoA) vanishes for f_A = 1/2
    assert np.max(np.abs(wabm)) < 1e-12


def test_1s_fixed_point_is_stationary():
    # the 1S map leaves w unchanged iff F + (A - B) w = 0; iterate a small
    # system to its saddle point and check H and w stop changing.
    w = _small(grid=(4, 4, 8), box_rg=(1.5, 1.5, 3.0), N=6)
    p = w.params()
    p.lam_imag = p.lam_real = 0.02   # reference units
    o = po.Oracle(p)
    wpl, wabp, wabm = seeded_fields(p, 11, period_rg=3.0, amp=0.3, sigma=0.01)
    o.set_fields(wpl, wabp, wabm)
    o.calc_poly_density()
    for _ in range(900):
        o.update_1s()
    H1 = o.calc_H()
    f1 = o.array("wabm")
    o.update_1s()
    H2 = o.calc_H()
    assert abs(H2 - H1) < 1e-9 * abs(H1)
    assert np.max(np.abs(o.array("wabm") - f1)) < 1e-8
