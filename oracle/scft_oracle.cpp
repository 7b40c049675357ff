/*
 * ORACLE -- test infrastructure only.
 *
 * A clean-room FP64 CPU restatement of the reference SCFT hot path
 * (rar-ensemble/aligned-nanoplate-separation-and-rotation, code/).  The HIP
 * library's results are compared against it: tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it; the product never links or calls
 * it.
 *
 * It follows the reference dataflow on purpose (it is also the timed CPU
 * baseline): forward FFT scaled by 1/M, a complex exp(-w) evaluated at every
 * propagator step, q and q-dagger stored as N separate arrays, and the
 * density sum over segments per grid point.  Pointwise loops and FFT lines
 * are threaded with OpenMP in place of the reference's MPI slabs; the grid
 * integral keeps the reference's serial index order.
 *
 * Pinning (tests/test_oracle.py):
 *   FFT            numpy.fft golden vectors, naive DFT, Parseval, plane waves
 *   kernels        closed forms, k=0 limits of the Debye functions
 *   cavity/volumes the LOG known answers of the published runs
 *   propagator/Q   w=0 -> Q=1; w=c -> Q=exp(-N c); plane-wave decay
 *   density        sum rule nD*N = integral(rhoA+rhoB), uniform-field values
 *   1S update      fixed point <=> vanishing force; single-mode response
 *   H              uniform-field closed form
 *   grafts, pairs  (tests/test_oracle_grafts.py) analytic graft-site
 *                  gradient, graft sum rule and pairing, uniform-field
 *                  Qga = -Ng c, force = dH/dw in every positive-chi field
 *   1S loop        through the library, which equals this oracle stage by
 *                  stage and reproduces the published runs' printed H at the
 *                  reference's stop rule from their own inputs
 *   Euler          (tests/test_oracle_euler.py) the noise-free step is
 *                  w - lambda dH/dw in every field of positive chi (w+, AB,
 *                  and with grafted chains BC and AC), with the reference's
 *                  lam_imag / lam_real assignment; the CL noise is pinned
 *                  statistically only, and no reference trajectory exists
 *                  (start fields missing).
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/scft.h"

typedef std::complex<double> cd;
static const double kPi = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------------ */
/* Mixed-radix Stockham FFT (own implementation; radices 2..7 and any other  */
/* prime through a direct small DFT).  Roots are taken from long double      */
/* sin/cos so each table entry is accurate to the last bit of a double.      */
/* ------------------------------------------------------------------------ */
struct Fft1 {
  int n = 1;
  std::vector<int> radix;
  std::vector<cd> root;  // root[m] = exp(-2 pi i m / n)

  explicit Fft1(int n_ = 1) : n(n_) {
    int m = n;
    const int pref[] = {4, 2, 3, 5, 7};
    for (int r : pref)
      while (m % r == 0) { radix.push_back(r); m /= r; }
    for (int p = 11; m > 1; p += 2)
      while (m % p == 0) { radix.push_back(p); m /= p; }
    root.resize(n);
    for (int k = 0; k < n; k++) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * k / n;
      root[k] = cd((double)cosl(a), (double)sinl(a));
    }
  }

  /* sign = -1: forward e^{-i}; +1: backward e^{+i}.  Unnormalised. */
  void run(cd* x, cd* scratch, int sign) const {
    if (n == 1) return;
    cd* a = x;
    cd* b = scratch;
    int Ns = 1;
    cd v[64], t[64];
    for (int R : radix) {
      const int stride = n / R;
      const int tw_step = n / (Ns * R);
      for (int j = 0; j < n / R; j++) {
        const int k = j % Ns;
        for (int r = 0; r < R; r++) {
          cd w = root[(size_t)(r * k * tw_step) % n];
          if (sign > 0) w = std::conj(w);
          v[r] = a[j + r * stride] * w;
        }
        // direct DFT of length R with roots of unity of order R
        for (int m = 0; m < R; m++) {
          cd s = 0.0;
          for (int r = 0; r < R; r++) {
            cd w = root[(size_t)((r * m) % R) * (n / R)];
            if (sign > 0) w = std::conj(w);
            s += v[r] * w;
          }
          t[m] = s;
        }
        const int base = (j / Ns) * Ns * R + k;
        for (int r = 0; r < R; r++) b[base + r * Ns] = t[r];
      }
      std::swap(a, b);
      Ns *= R;
    }
    if (a != x) std::memcpy(x, a, sizeof(cd) * n);
  }
};

/* 3D transform over the reference layout x + Nx*(y + Ny*z). */
struct Fft3 {
  int nx, ny, nz;
  Fft1 fx, fy, fz;
  Fft3(int a, int b, int c) : nx(a), ny(b), nz(c), fx(a), fy(b), fz(c) {}

  void run(cd* d, int sign) const {
    const long M = (long)nx * ny * nz;
    (void)M;
    if (nx > 1) {
#pragma omp parallel
      {
        std::vector<cd> s(nx);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)ny * nz; l++) fx.run(d + l * nx, s.data(), sign);
      }
    }
    if (ny > 1) {
#pragma omp parallel
      {
        std::vector<cd> line(ny), s(ny);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)nx * nz; l++) {
          const long x = l % nx, z = l / nx;
          cd* base = d + x + (long)nx * ny * z;
          for (int i = 0; i < ny; i++) line[i] = base[(long)i * nx];
          fy.run(line.data(), s.data(), sign);
          for (int i = 0; i < ny; i++) base[(long)i * nx] = line[i];
        }
      }
    }
    if (nz > 1) {
#pragma omp parallel
      {
        std::vector<cd> line(nz), s(nz);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)nx * ny; l++) {
          cd* base = d + l;
          const long st = (long)nx * ny;
          for (int i = 0; i < nz; i++) line[i] = base[i * st];
          fz.run(line.data(), s.data(), sign);
          for (int i = 0; i < nz; i++) base[i * st] = line[i];
        }
      }
    }
  }
};

/* ------------------------------------------------------------------------ */
/* Numerical Recipes ran2 (L'Ecuyer two-generator + Bays-Durham shuffle) and */
/* the polar Box-Muller gasdev, restated from the published algorithm that   */
/* code/ran2.cpp:32-90 uses.                                                  */
/* ------------------------------------------------------------------------ */
struct Ran2 {
  long idum = -1, idum2 = 123456789, iy = 0;
  long iv[32];
  int have_g = 0;
  double g_saved = 0.0;

  double uniform() {
    const long m1 = 2147483563, m2 = 2147483399, a1 = 40014, a2 = 40692;
    const long q1 = 53668, q2 = 52774, r1 = 12211, r2 = 3791;
    const long ntab = 32, mm1 = m1 - 1, ndiv = 1 + mm1 / ntab;
    const double scale = 1.0 / m1, top = 1.0 - 1.2e-7;
    if (idum <= 0) {
      idum = (-idum < 1) ? 1 : -idum;
      idum2 = idum;
      for (int j = ntab + 7; j >= 0; j--) {
        long h = idum / q1;
        idum = a1 * (idum - h * q1) - h * r1;
        if (idum < 0) idum += m1;
        if (j < ntab) iv[j] = idum;
      }
      iy = iv[0];
    }
    long h = idum / q1;
    idum = a1 * (idum - h * q1) - h * r1;
    if (idum < 0) idum += m1;
    h = idum2 / q2;
    idum2 = a2 * (idum2 - h * q2) - h * r2;
    if (idum2 < 0) idum2 += m2;
    int j = (int)(iy / ndiv);
    iy = iv[j] - idum2;
    iv[j] = idum;
    if (iy < 1) iy += mm1;
    double t = scale * iy;
    return t > top ? top : t;
  }

  double gauss() {
    if (have_g) { have_g = 0; return g_saved; }
    double v1, v2, rsq;
    do {
      v1 = 2.0 * uniform() - 1.0;
      v2 = 2.0 * uniform() - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (rsq >= 1.0 || rsq == 0.0);
    double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
    g_saved = v1 * fac;
    have_g = 1;
    return v2 * fac;
  }
};

/* ------------------------------------------------------------------------ */
/* The system state (the reference keeps all of it in globals).              */
/* ------------------------------------------------------------------------ */
struct Oracle {
  scft_params p;
  int Dim;
  int Nx[3];
  double L[3], dx[3];
  long M;
  double V;
  int N, Nda, Ndb;
  double fD, Rg2, a_sq;
  double chi_ab, chi_bc, chi_ac;
  int neg_ab, neg_bc, neg_ac;
  double Vf = 0, V_1_exp_np = 0, np_frac = 0, phiD = 0, phiG = 0, nP = 0,
         ng_per_np = 0, nD = 0;
  Fft3* fft = nullptr;
  std::vector<cd> buf;  // staging copy of the FFT wrapper
  std::vector<cd> wpl, wabp, wabm, wbcp, wbcm, wacp, wacm;
  std::vector<cd> wa, wb, wg, smwa, smwb, smwg;
  std::vector<cd> rhoda, rhodb, rho_exp_nr, exp_nrH, rho_surf, surfH;
  std::vector<cd> hhat, bond, gaa, gab, gbb, gc;
  std::vector<cd> etap, etam, tmp;
  std::vector<std::vector<cd>> q, qdag;
  // grafted chains on the explicit plates: graft-site density, chain
  // propagators, graft density and the log-average Qga_exp
  std::vector<cd> expl_grafts, rhog_exp;
  std::vector<std::vector<cd>> qg, qgdag;
  cd Qd = 1.0, Hcur = 0.0, Qga = 0.0;
  Ran2 rng;
  int iter = 0;

  /* index helpers, code/array_utils.cpp:80-87, 137-157 */
  void unstack(long id, int n[3]) const {
    n[2] = (int)(id / ((long)Nx[0] * Nx[1]));
    n[1] = (int)(id / Nx[0] - (long)n[2] * Nx[1]);
    n[0] = (int)(id - ((long)n[1] + (long)n[2] * Nx[1]) * Nx[0]);
  }
  /* get_k, code/array_utils.cpp:239-272: n < Nx/2 -> +, else n - Nx */
  double get_k(long id, double k[3]) const {
    int n[3];
    unstack(id, n);
    double k2 = 0.0;
    for (int d = 0; d < 3; d++) {
      if (d >= Dim) { k[d] = 0.0; continue; }
      if (double(n[d]) < double(Nx[d]) / 2.0)
        k[d] = 2 * kPi * double(n[d]) / L[d];
      else
        k[d] = 2 * kPi * double(n[d] - Nx[d]) / L[d];
      k2 += k[d] * k[d];
    }
    return k2;
  }
  /* pbc_mdr2, code/array_utils.cpp:588-601 */
  double pbc_mdr2(const double* x1, const double* x2, double* dr) const {
    double s = 0.0;
    for (int d = 0; d < Dim; d++) {
      dr[d] = x1[d] - x2[d];
      if (dr[d] >= 0.5 * L[d]) dr[d] -= L[d];
      else if (dr[d] < -0.5 * L[d]) dr[d] += L[d];
      s += dr[d] * dr[d];
    }
    return s;
  }
  double dot(const double* u, const double* r) const {
    double s = 0;
    for (int d = 0; d < Dim; d++) s += u[d] * r[d];
    return s;
  }
  /* cross_prod, code/array_utils.cpp:629-647 */
  double cross(const double* u, const double* r) const {
    if (Dim == 2) return std::fabs(u[0] * r[1] - u[1] * r[0]);
    double c0 = u[1] * r[2] - u[2] * r[1];
    double c1 = u[2] * r[0] - u[0] * r[2];
    double c2 = u[0] * r[1] - u[1] * r[0];
    return std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }

  /* integ_trapPBC, code/integrate_utils.cpp:24-45 (serial index order) */
  cd integ(const std::vector<cd>& a) const {
    cd s = 0.0;
    for (long i = 0; i < M; i++) s += a[i];
    for (int d = 0; d < Dim; d++) s *= dx[d];
    return s;
  }

  /* fft_fwd_wrapper / fft_bck_wrapper, code/fft_wrappers.cpp:4-67: copy into
   * a staging buffer, transform, copy out (forward scaled by 1/M). */
  void fft_fwd(const cd* in, cd* out) {
    std::memcpy(buf.data(), in, sizeof(cd) * M);
    fft->run(buf.data(), -1);
    const double norm = 1.0 / double(M);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) out[i] = buf[i] * norm;
  }
  void fft_bck(const cd* in, cd* out) {
    std::memcpy(buf.data(), in, sizeof(cd) * M);
    fft->run(buf.data(), +1);
    std::memcpy(out, buf.data(), sizeof(cd) * M);
  }
  void fft_fwd(std::vector<cd>& a) { fft_fwd(a.data(), a.data()); }
  void fft_bck(std::vector<cd>& a) { fft_bck(a.data(), a.data()); }

  /* Plate density.  np_type 2: explicit_nanorod, code/initialize.cpp:578-612,
   * rho = 1/4 erfc((|u.dr|-len/2)/xi) erfc((|u x dr|-rad)/xi), with x = n dx
   * and the minimum image of pbc_mdr2.  np_type 3 (extension): square plate of
   * half side rad, 1/8 erfc((|u.dr|-len/2)/xi) erfc((|e1.dr|-rad)/xi)
   * erfc((|e2.dr|-rad)/xi); in 2D it reduces to np_type 2. */
  void plate_density(double len, double rad, double xi, const double* rel_c,
                     const double* u, std::vector<cd>& out) const {
    double c[3] = {0, 0, 0};
    for (int d = 0; d < Dim; d++) c[d] = rel_c[d] * L[d];
    double e1[3], e2[3];
    plate_axes(u, e1, e2);
    const int square = (p.np_type == SCFT_NP_SQUARE && Dim == 3);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      int n[3];
      unstack(i, n);
      double x[3] = {0, 0, 0}, dr[3] = {0, 0, 0};
      for (int d = 0; d < Dim; d++) x[d] = double(n[d]) * dx[d];
      pbc_mdr2(x, c, dr);
      double ud = dot(u, dr);
      if (!square) {
        double uc = cross(u, dr);
        out[i] = 0.25 * std::erfc((std::fabs(ud) - 0.5 * len) / xi) *
                 std::erfc((std::fabs(uc) - rad) / xi);
      } else {
        double a1 = dot(e1, dr), a2 = dot(e2, dr);
        out[i] = 0.125 * std::erfc((std::fabs(ud) - 0.5 * len) / xi) *
                 std::erfc((std::fabs(a1) - rad) / xi) *
                 std::erfc((std::fabs(a2) - rad) / xi);
      }
    }
  }
  /* In-plane axes of a square plate with unit normal u: e2 = u x X (or u x Y
   * when u is along X), e1 = e2 x u. */
  static void plate_axes(const double* u, double* e1, double* e2) {
    double ax[3] = {1, 0, 0};
    double c[3] = {u[1] * ax[2] - u[2] * ax[1], u[2] * ax[0] - u[0] * ax[2],
                   u[0] * ax[1] - u[1] * ax[0]};
    double n = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (n < 1e-12) {
      double ay[3] = {0, 1, 0};
      c[0] = u[1] * ay[2] - u[2] * ay[1];
      c[1] = u[2] * ay[0] - u[0] * ay[2];
      c[2] = u[0] * ay[1] - u[1] * ay[0];
      n = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    }
    for (int d = 0; d < 3; d++) e2[d] = c[d] / n;
    e1[0] = e2[1] * u[2] - e2[2] * u[1];
    e1[1] = e2[2] * u[0] - e2[0] * u[2];
    e1[2] = e2[0] * u[1] - e2[1] * u[0];
  }

  /* initialize_1, code/initialize.cpp:45-81 */
  void initialize_1() {
    Dim = p.dim;
    Nda = p.Nda;
    Ndb = p.Ndb;
    N = Nda + Ndb;
    fD = N > 0 ? double(Nda) / double(N) : 0.0;
    Rg2 = double(N - 1) / 6.0;
    a_sq = p.a_smear * p.a_smear;
    M = 1;
    V = 1.0;
    for (int d = 0; d < 3; d++) {
      Nx[d] = d < Dim ? p.Nx[d] : 1;
      L[d] = d < Dim ? p.L[d] : 1.0;
      dx[d] = L[d] / double(Nx[d]);
      M *= Nx[d];
      if (d < Dim) V *= L[d];
    }
    neg_ab = p.chi_ab < 0;
    chi_ab = std::fabs(p.chi_ab);
    neg_bc = p.chi_bc < 0;
    chi_bc = std::fabs(p.chi_bc);
    neg_ac = p.chi_ac < 0;
    chi_ac = std::fabs(p.chi_ac);
  }

  void allocate() {
    auto z = [&](std::vector<cd>& a) { a.assign(M, cd(0.0, 0.0)); };
    for (auto* a : {&buf, &wpl, &wabp, &wabm, &wbcp, &wbcm, &wacp, &wacm, &wa,
                    &wb, &wg, &smwa, &smwb, &smwg, &rhoda, &rhodb, &rho_exp_nr,
                    &exp_nrH, &rho_surf, &surfH, &hhat, &bond, &gaa, &gab, &gbb, &gc,
                    &etap, &etam, &tmp})
      z(*a);
    q.assign(N + 1, std::vector<cd>(M));
    qdag.assign(N + 1, std::vector<cd>(M));
    z(expl_grafts);
    z(rhog_exp);
    if (grafted()) {
      qg.assign(p.Ng, std::vector<cd>(M));
      qgdag.assign(p.Ng, std::vector<cd>(M));
    }
    fft = new Fft3(Nx[0], Nx[1], Nx[2]);
  }

  /* grafted chains on explicit plates (sigma > 0, Ng > 0, rods) */
  bool grafted() const { return p.n_exp_nr > 0 && p.sigma > 0.0 && p.Ng > 0; }

  /* get_k_alias, code/array_utils.cpp:181-230: the Nyquist mode of an even
   * axis gets k = 0 */
  void get_k_alias(long id, double k[3]) const {
    int n[3];
    unstack(id, n);
    for (int d = 0; d < 3; d++) {
      if (d >= Dim) { k[d] = 0.0; continue; }
      if (Nx[d] % 2 == 0 && n[d] == Nx[d] / 2) k[d] = 0.0;
      else if (double(n[d]) < double(Nx[d]) / 2.0) k[d] = 2 * kPi * double(n[d]) / L[d];
      else k[d] = 2 * kPi * double(n[d] - Nx[d]) / L[d];
    }
  }
  /* field_gradient, code/array_utils.cpp:565-579: spectral derivative */
  void field_gradient(const std::vector<cd>& in, std::vector<cd>& out, int dir) {
    const cd I(0.0, 1.0);
    fft_fwd(in.data(), out.data());
    for (long i = 0; i < M; i++) {
      double kv[3];
      get_k_alias(i, kv);
      out[i] *= I * kv[dir];
    }
    fft_bck(out);
  }

  /* graft sites, code/initialize.cpp:160-196: |grad| of each plate's shape
   * inflated by 1.05 in length and radius, summed, normalised to unit
   * integral */
  void graft_sites() {
    std::vector<cd> shape(M), g(M), g2(M);
    for (int np = 0; np < p.n_exp_nr; np++) {
      plate_density(p.L_nr * 1.05, p.R_nr * 1.05, p.xi_nr, p.exp_nr_c[np], p.exp_nr_u[np], shape);
      std::fill(g2.begin(), g2.end(), cd(0.0, 0.0));
      for (int d = 0; d < Dim; d++) {
        field_gradient(shape, g, d);
        for (long i = 0; i < M; i++) g2[i] += g[i] * g[i];
      }
      for (long i = 0; i < M; i++) expl_grafts[i] += std::sqrt(g2[i]);
    }
    const cd total = integ(expl_grafts);
    for (long i = 0; i < M; i++) expl_grafts[i] /= total;
  }

  /* initialize_2, code/initialize.cpp:86-412 (film, channel and field NPs are
   * not on this path). */
  void initialize_2() {
    double k[3];
    for (long i = 0; i < M; i++) hhat[i] = std::exp(-get_k(i, k) * a_sq / 2.0);
    if (grafted()) graft_sites();
    // plates, :210-239
    if (p.n_exp_nr > 0) {
      std::vector<cd> one(M);
      for (int np = 0; np < p.n_exp_nr; np++) {
        plate_density(p.L_nr, p.R_nr, p.xi_nr, p.exp_nr_c[np], p.exp_nr_u[np], one);
        for (long i = 0; i < M; i++) rho_exp_nr[i] += one[i];
      }
      fft_fwd(rho_exp_nr.data(), exp_nrH.data());
    }
    init_fields();
    // volumes, :249-355
    double V_exp = integ(rho_exp_nr).real();
    Vf = V - integ(rho_surf).real();
    double V_nps = V_exp;
    np_frac = V_nps / Vf;
    V_1_exp_np = p.n_exp_nr > 0 ? V_exp / double(p.n_exp_nr) : 0.0;
    nP = (p.n_exp_nr == 0) ? 0.0 : V_nps / V_1_exp_np;
    if (p.sigma > 0.0) {
      ng_per_np = (Dim == 2) ? p.sigma * 2.0 * kPi * p.R_nr
                             : p.sigma * 4.0 * kPi * p.R_nr * p.R_nr;
      phiG = ng_per_np * nP * p.Ng / p.rho0 / Vf;
    } else {
      ng_per_np = 0.0;
      phiG = 0.0;
    }
    phiD = 1.0 - p.phiH - np_frac - phiG;
    nD = phiD * p.rho0 * Vf / N;
    // Debye functions, code/calc_debye.cpp:5-51
    for (long i = 0; i < M; i++) {
      double x = get_k(i, k) * Rg2;
      gaa[i] = 2.0 * (std::exp(-x * fD) + fD * x - 1.0) / x / x;
      gbb[i] = 2.0 * (std::exp(-x * (1.0 - fD)) + (1.0 - fD) * x - 1.0) / x / x;
      gab[i] = (1.0 - std::exp(-x * fD)) * (1.0 - std::exp(-x * (1.0 - fD))) / x / x;
    }
    gaa[0] = fD * fD;
    gbb[0] = (1.0 - fD) * (1.0 - fD);
    gab[0] = fD * (1.0 - fD);
    // graft Debye function, calc_gd (code/calc_debye.cpp:53-68) with
    // alpha = (Ng-1)/(N-1), code/initialize.cpp:394
    if (p.sigma > 0) {
      const double alpha = double(p.Ng - 1) / double(N - 1);
      for (long i = 0; i < M; i++) {
        double d = get_k(i, k) * Rg2 * alpha;
        gc[i] = (d == 0.0) ? 1.0 : 2.0 * (std::exp(-d) + d - 1.0) / d / d;
      }
    }
    // Gaussian bond, :401-404
    for (long i = 0; i < M; i++) bond[i] = std::exp(-get_k(i, k) / 6.0);
    iter = 0;
  }

  /* init_fields, code/initialize.cpp:419-565 (before the .res override). */
  void init_fields() {
    const cd I(0.0, 1.0);
    for (long i = 0; i < M; i++) {
      wpl[i] = wabp[i] = wabm[i] = wacp[i] = wacm[i] = wbcp[i] = wbcm[i] = 0.0;
      int n[3];
      unstack(i, n);
      double x[3];
      for (int d = 0; d < 3; d++) {
        x[d] = dx[d] * n[d];
        if (x[d] > L[d] / 2.0) x[d] -= L[d];
        else if (x[d] <= -L[d] / 2.0) x[d] += L[d];
      }
      if (p.ic_flag[0] < -1) {
        if ((rho_surf[i] + rho_exp_nr[i]).real() > 0.5) {
          wabm[i] = 0.0;
        } else {
          cd v = 1.0;
          for (int j = 0; j < 2; j++) {
            int dir = int(p.ic_dir[j]);
            cd ph = (chi_ab != 0 ? (neg_ab ? I : cd(1.0)) : cd(1.0));
            if (p.ic_flag[j] == -2)
              v *= p.ic_pre[j] * std::tanh(std::cos(2.0 * kPi * p.ic_period[j] * x[dir] / L[dir]) / 0.2) * ph;
            else if (p.ic_flag[j] == -3)
              v *= p.ic_pre[j] * std::tanh(std::sin(2.0 * kPi * p.ic_period[j] * x[dir] / L[dir]) / 0.2) * ph;
          }
          wabm[i] = v;
        }
      } else if (p.ic_flag[0] == -1) {
        if (chi_ab != 0) {
          wabp[i] = p.ic_pre[0] * (rng.uniform() - 0.5) * (neg_ab ? cd(1.0) : I);
          wabm[i] = p.ic_pre[0] * (rng.uniform() - 0.5) * (neg_ab ? I : cd(1.0));
        }
        if (chi_bc != 0) {
          wbcp[i] = p.ic_pre[0] * (rng.uniform() - 0.5) * (neg_bc ? cd(1.0) : I);
          wbcm[i] = p.ic_pre[0] * (rng.uniform() - 0.5) * (neg_bc ? I : cd(1.0));
        }
        if (chi_ac != 0) {
          wacp[i] = p.ic_pre[0] * (rng.uniform() - 0.5) * (neg_ac ? cd(1.0) : I);
          wacm[i] = p.ic_pre[0] * (rng.uniform() - 0.5) * (neg_ac ? I : cd(1.0));
        }
        wpl[i] = p.ic_pre[0] * (rng.uniform() - 0.5) * I;
      }
    }
  }

  /* diblock_discrete, code/diblock_discrete.cpp:4-95 */
  cd diblock_discrete(const std::vector<cd>& WA, const std::vector<cd>& WB) {
    std::vector<cd> a(M), h(M), qf(M);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      q[0][i] = (Nda != 0) ? std::exp(-WA[i]) : std::exp(-WB[i]);
      qdag[0][i] = (Nda != N) ? std::exp(-WB[i]) : std::exp(-WA[i]);
    }
    for (int n = 1; n < N; n++) {
      const std::vector<cd>& w = (Nda > 0 && n < Nda) ? WA : WB;
      fft_fwd(q[n - 1].data(), a.data());
#pragma omp parallel for schedule(static)
      for (long j = 0; j < M; j++) h[j] = a[j] * bond[j];
      fft_bck(h.data(), qf.data());
#pragma omp parallel for schedule(static)
      for (long j = 0; j < M; j++) q[n][j] = qf[j] * std::exp(-w[j]);
      if (Nda != N && Nda != 0) {
        const std::vector<cd>& wd = (n <= N - Nda - 1) ? WB : WA;
        fft_fwd(qdag[n - 1].data(), a.data());
#pragma omp parallel for schedule(static)
        for (long j = 0; j < M; j++) h[j] = a[j] * bond[j];
        fft_bck(h.data(), qf.data());
#pragma omp parallel for schedule(static)
        for (long j = 0; j < M; j++) qdag[n][j] = qf[j] * std::exp(-wd[j]);
      } else {
        qdag[n] = q[n];
      }
    }
    return integ(q[N - 1]) / V;
  }

  /* integrate_diblock_discrete, code/calc_poly_density.cpp:167-191 */
  void integrate_diblock(cd Q) {
    const cd factor = cd(nD) / Q / V;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      cd ra = 0.0, rb = 0.0;
      for (int j = 0; j < N; j++) {
        if (Nda > 0 && j < Nda) ra += q[j][i] * qdag[N - j - 1][i];
        else rb += q[j][i] * qdag[N - j - 1][i];
      }
      rhoda[i] = ra * std::exp(smwa[i]) * factor;
      rhodb[i] = rb * std::exp(smwb[i]) * factor;
    }
  }

  /* calc_poly_density, code/calc_poly_density.cpp:37-122 (diblock part; the
   * smeared graft field wg is formed and smeared as there). */
  void calc_poly_density() {
    const cd I(0.0, 1.0);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      wa[i] = I * wpl[i] + (neg_ab ? I * wabm[i] - wabp[i] : I * wabp[i] - wabm[i]) +
              (neg_ac ? I * wacm[i] - wacp[i] : I * wacp[i] - wacm[i]);
      wb[i] = I * wpl[i] + (neg_ab ? -I * wabm[i] - wabp[i] : I * wabp[i] + wabm[i]) +
              (neg_bc ? I * wbcm[i] - wbcp[i] : I * wbcp[i] - wbcm[i]);
      wg[i] = I * wpl[i] + (neg_ac ? -I * wacm[i] - wacp[i] : I * wacp[i] + wacm[i]) +
              (neg_bc ? -I * wbcm[i] - wbcp[i] : I * wbcp[i] + wbcm[i]);
    }
    fft_fwd(wa.data(), smwa.data());
    fft_fwd(wb.data(), smwb.data());
    fft_fwd(wg.data(), smwg.data());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      smwa[i] *= hhat[i];
      smwb[i] *= hhat[i];
      smwg[i] *= hhat[i];
    }
    fft_bck(smwa);
    fft_bck(smwb);
    fft_bck(smwg);
    if (nD > 0.0) {
      Qd = diblock_discrete(smwa, smwb);
      integrate_diblock(Qd);
    } else {
      Qd = 1.0;
    }
    if (grafted()) {
      graft_free_ends();
      Qga = grafted_exp_nps();
    }
  }

  /* one step of a chain: out = exp(-w) IFFT[bond FFT[in]],
   * code/grafted_nanoparticle.cpp:29-38 */
  void chain_step(const std::vector<cd>& in, std::vector<cd>& out, const std::vector<cd>& w) {
    fft_fwd(in.data(), out.data());
    for (long i = 0; i < M; i++) out[i] *= bond[i];
    fft_bck(out);
    for (long i = 0; i < M; i++) out[i] *= std::exp(-w[i]);
  }
  /* graft_homopoly_free_ends, code/grafted_nanoparticle.cpp:15-41 */
  void graft_free_ends() {
    for (long i = 0; i < M; i++) qg[0][i] = std::exp(-smwg[i]);
    for (int j = 1; j < p.Ng; j++) chain_step(qg[j - 1], qg[j], smwg);
  }
  /* grafted_exp_nps, code/grafted_nanoparticle.cpp:43-117: the complementary
   * propagator starts at the graft sites, divided by the free-end
   * propagator; returns Qga_exp = integ(expl_grafts log q_g[Ng-1]) */
  cd grafted_exp_nps() {
    const int Ng = p.Ng;
    for (long i = 0; i < M; i++) qgdag[0][i] = expl_grafts[i] * std::exp(-smwg[i]) / qg[Ng - 1][i];
    for (int j = 1; j < Ng; j++) chain_step(qgdag[j - 1], qgdag[j], smwg);
    for (long i = 0; i < M; i++) {
      cd r = 0.0;
      for (int j = 0; j < Ng; j++) r += qg[j][i] * qgdag[Ng - j - 1][i];
      rhog_exp[i] = r * (ng_per_np * double(p.n_exp_nr) * std::exp(smwg[i]));
    }
    for (long i = 0; i < M; i++)
      tmp[i] = (expl_grafts[i].real() <= 0.0) ? cd(0.0) : expl_grafts[i] * std::log(qg[Ng - 1][i]);
    return integ(tmp);
  }

  /* generate_1s_noise, code/1s_update.cpp:356-368 */
  void noise(std::vector<cd>& et, double lambda) {
    double scale = std::sqrt(2.0 * lambda);
    for (int d = 0; d < Dim; d++) scale /= std::sqrt(dx[d]);
    for (long i = 0; i < M; i++) et[i] = scale * rng.gauss();
    fft_fwd(et);
  }

  /* One graft pair of update_1s: BC (rho = rhodb, g = gbb,
   * code/1s_update.cpp:162-247) or AC (rho = rhoda, g = gaa, :249-329), for
   * both signs of chi; the densities are in k-space. */
  void pair_update(std::vector<cd>& wp, std::vector<cd>& wm, const std::vector<cd>& rho,
                   const std::vector<cd>& g, double chi, bool neg) {
    const cd I(0.0, 1.0);
    const double rho0 = p.rho0, li = p.lam_imag, lr = p.lam_real;
    const int Ng = p.Ng;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      if (!neg) {
        cd F = I * hhat[i] * (rhog_exp[i] + rho[i]);
        cd A = 2.0 * rho0 / chi + phiD * rho0 * N * hhat[i] * hhat[i] * (g[i]) +
               phiG * rho0 * Ng * hhat[i] * hhat[i] * gc[i];
        cd B = A - 2.0 * rho0 / chi;
        cd numer = wp[i] - li * (F - B * wp[i]);
        if (p.do_CL) numer += etap[i];
        wp[i] = numer / (1.0 + li * A);
        F = hhat[i] * (rhog_exp[i] - rho[i]);
        if (p.np_chem == 1) F += I * rho0 * exp_nrH[i];
        numer = wm[i] - lr * F;
        if (p.do_CL) numer += etam[i];
        wm[i] = numer / (1.0 + lr * 2.0 * rho0 / chi);
      } else {
        cd F = -hhat[i] * (rho[i] + rhog_exp[i]);
        cd numer = wp[i] - lr * F;
        if (p.do_CL) numer += etap[i];
        wp[i] = numer / (1.0 + lr * 2.0 * rho0 / chi);
        F = -I * hhat[i] * (rhog_exp[i] - rho[i]);
        if (p.np_chem == 1) F += -rho0 * exp_nrH[i];
        cd A = 2.0 * rho0 / chi + phiD * rho0 * N * hhat[i] * hhat[i] * (g[i]) -
               phiG * Ng * hhat[i] * hhat[i] * rho0 * gc[i];
        cd B = A - 2.0 * rho0 / chi;
        numer = wm[i] - li * (F - B * wm[i]);
        if (p.do_CL) numer += etap[i];
        wm[i] = numer / (1.0 + li * A);
      }
    }
  }

  /* update_1s, code/1s_update.cpp:7-352 (homopolymer, field-NP and channel
   * densities are zero on this path). */
  void update_1s() {
    const cd I(0.0, 1.0);
    const double rho0 = p.rho0, kappa = p.kappa, li = p.lam_imag, lr = p.lam_real;
    if (nD > 0.0) {
      fft_fwd(rhoda);
      fft_fwd(rhodb);
    }
    if (p.sigma > 0.0 && p.n_exp_nr > 0) fft_fwd(rhog_exp);
    fft_fwd(wpl);
    if (chi_ab != 0.0) { fft_fwd(wabm); fft_fwd(wabp); }
    if (chi_bc != 0.0) { fft_fwd(wbcm); fft_fwd(wbcp); }
    if (chi_ac != 0.0) { fft_fwd(wacm); fft_fwd(wacp); }
    if (p.do_CL) noise(etap, li);
    // w+ (:57-84)
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      cd evar = (i == 0) ? 1.0 : 0.0;
      cd F = I * rho0 * (surfH[i] + exp_nrH[i] - evar) + I * hhat[i] * (rhoda[i] + rhog_exp[i] + rhodb[i]);
      cd A = (kappa <= 0.0 ? 0.0 : rho0 / kappa) +
             phiD * rho0 * N * hhat[i] * hhat[i] * (gaa[i] + 2.0 * gab[i] + gbb[i]) +
             (phiG * p.Ng * gc[i]) * rho0 * hhat[i] * hhat[i];
      cd B = kappa > 0 ? A - rho0 / kappa : cd(0.0);
      cd numer = wpl[i] - li * (F - B * wpl[i]);
      if (p.do_CL) numer += etap[i];
      wpl[i] = numer / (1.0 + li * A);
    }
    if (chi_ab != 0.0) {
      if (p.do_CL) { noise(etap, lr); noise(etam, li); }
#pragma omp parallel for schedule(static)
      for (long i = 0; i < M; i++) {
        if (!neg_ab) {
          cd F = I * hhat[i] * (rhoda[i] + rhodb[i]);
          if (p.np_chem == 1) F += -rho0 * exp_nrH[i];
          cd A = 2.0 * rho0 / chi_ab + phiD * rho0 * N * hhat[i] * hhat[i] * (gaa[i] + 2.0 * gab[i] + gbb[i]);
          cd B = A - 2.0 * rho0 / chi_ab;
          cd numer = wabp[i] - li * (F - B * wabp[i]);
          if (p.do_CL) numer += etap[i];
          wabp[i] = numer / (1.0 + li * A);
          F = hhat[i] * (rhodb[i] - rhoda[i]);
          if (p.np_chem == 1) F += I * rho0 * exp_nrH[i];
          numer = wabm[i] - lr * F;
          if (p.do_CL) numer += etam[i];
          wabm[i] = numer / (1.0 + lr * 2.0 * rho0 / chi_ab);
        } else {
          cd F = -hhat[i] * (rhoda[i] + rhodb[i]);
          wabp[i] = (wabp[i] - lr * F) / (1.0 + lr * 2.0 * rho0 / chi_ab);
          F = I * hhat[i] * (rhoda[i] - rhodb[i]);
          if (p.np_chem == 1) F += -rho0 * exp_nrH[i];
          cd A = 2.0 * rho0 / chi_ab + phiD * rho0 * N * hhat[i] * hhat[i] * (gaa[i] - gbb[i]);
          cd B = A - 2.0 * rho0 / chi_ab;
          cd numer = wabm[i] - li * (F - B * wabm[i]);
          if (p.do_CL) numer += etap[i];
          wabm[i] = numer / (1.0 + li * A);
        }
      }
    } else {
      for (long i = 0; i < M; i++) wabp[i] = wabm[i] = 0.0;
    }
    // BC and AC pairs (:162-329): the graft chains against the B and A blocks
    if (chi_bc != 0.0) {
      if (p.do_CL) { noise(etap, lr); noise(etam, li); }
      pair_update(wbcp, wbcm, rhodb, gbb, chi_bc, neg_bc);
    } else {
      for (long i = 0; i < M; i++) wbcp[i] = wbcm[i] = 0.0;
    }
    if (chi_ac != 0.0) {
      if (p.do_CL) { noise(etap, lr); noise(etam, li); }
      pair_update(wacp, wacm, rhoda, gaa, chi_ac, neg_ac);
    } else {
      for (long i = 0; i < M; i++) wacp[i] = wacm[i] = 0.0;
    }
    fft_bck(wpl);
    if (chi_ab != 0.0) { fft_bck(wabm); fft_bck(wabp); }
    if (chi_bc != 0.0) { fft_bck(wbcm); fft_bck(wbcp); }
    if (chi_ac != 0.0) { fft_bck(wacm); fft_bck(wacp); }
    calc_poly_density();
  }

  /* One graft pair of update_Euler: BC (rho = rhodb,
   * code/Euler_update.cpp:108-143) or AC (rho = rhoda, :146-181), explicit
   * steps, the + field with lam_imag and the - field with lam_real
   * (:122,133,160,171); densities in k-space. */
  void euler_pair(std::vector<cd>& wp, std::vector<cd>& wm, const std::vector<cd>& rho, double chi, bool neg) {
    const cd I(0.0, 1.0);
    const double rho0 = p.rho0, li = p.lam_imag, lr = p.lam_real;
    const cd sp = neg ? cd(-1.0) : I;   // rho and rhog_exp in the + force
    const cd sm = neg ? I : cd(-1.0);   // rho in the - force
    const cd sg = neg ? -I : cd(1.0);   // rhog_exp in the - force
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      cd F = 2.0 * rho0 * wp[i] / chi + hhat[i] * (rho[i] * sp + rhog_exp[i] * sp);
      cd numer = wp[i] - li * F;
      if (p.do_CL) numer += etap[i];
      wp[i] = numer;
      F = 2.0 * rho0 * wm[i] / chi + hhat[i] * (rho[i] * sm + rhog_exp[i] * sg);
      numer = wm[i] - lr * F;
      if (p.do_CL) numer += etam[i];
      wm[i] = numer;
    }
  }

  /* update_Euler, code/Euler_update.cpp:5-213: w+, the AB pair and the BC /
   * AC pairs of the grafted chains. */
  void update_euler() {
    const cd I(0.0, 1.0);
    const double rho0 = p.rho0, kappa = p.kappa, li = p.lam_imag, lr = p.lam_real;
    if (nD > 0.0) { fft_fwd(rhoda); fft_fwd(rhodb); }
    if (p.sigma > 0.0 && p.n_exp_nr > 0) fft_fwd(rhog_exp);
    fft_fwd(wpl);
    if (chi_ab != 0.0) { fft_fwd(wabm); fft_fwd(wabp); }
    if (chi_bc != 0.0) { fft_fwd(wbcm); fft_fwd(wbcp); }
    if (chi_ac != 0.0) { fft_fwd(wacm); fft_fwd(wacp); }
    if (p.do_CL) noise(etap, li);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++) {
      cd evar = (i == 0) ? 1.0 : 0.0;
      cd F = (kappa <= 0.0 ? cd(0.0) : rho0 * wpl[i] / kappa) +
             I * rho0 * (surfH[i] + exp_nrH[i] - evar) + I * hhat[i] * (rhoda[i] + rhodb[i] + rhog_exp[i]);
      cd numer = wpl[i] - li * F;
      if (p.do_CL) numer += etap[i];
      wpl[i] = numer;
    }
    if (chi_ab != 0.0) {
      if (p.do_CL) { noise(etap, lr); noise(etam, li); }
      const cd sp = neg_ab ? cd(-1.0) : I;
      const cd sa = neg_ab ? I : cd(-1.0);
      const cd sb = neg_ab ? -I : cd(1.0);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < M; i++) {
        cd F = 2.0 * rho0 * wabp[i] / chi_ab + hhat[i] * (rhoda[i] * sp + rhodb[i] * sp);
        cd numer = wabp[i] - li * F;
        if (p.do_CL) numer += etap[i];
        wabp[i] = numer;
        F = 2.0 * rho0 * wabm[i] / chi_ab + hhat[i] * (rhoda[i] * sa + rhodb[i] * sb);
        numer = wabm[i] - lr * F;
        if (p.do_CL) numer += etam[i];
        wabm[i] = numer;
      }
    } else {
      for (long i = 0; i < M; i++) wabp[i] = wabm[i] = 0.0;
    }
    if (chi_bc != 0.0) {
      if (p.do_CL) { noise(etap, li); noise(etam, lr); }
      euler_pair(wbcp, wbcm, rhodb, chi_bc, neg_bc);
    } else {
      for (long i = 0; i < M; i++) wbcp[i] = wbcm[i] = 0.0;
    }
    if (chi_ac != 0.0) {
      if (p.do_CL) { noise(etap, lr); noise(etam, li); }
      euler_pair(wacp, wacm, rhoda, chi_ac, neg_ac);
    } else {
      for (long i = 0; i < M; i++) wacp[i] = wacm[i] = 0.0;
    }
    fft_bck(wpl);
    if (chi_ab != 0.0) { fft_bck(wabm); fft_bck(wabp); }
    if (chi_bc != 0.0) { fft_bck(wbcm); fft_bck(wbcp); }
    if (chi_ac != 0.0) { fft_bck(wacm); fft_bck(wacp); }
    calc_poly_density();
  }

  /* calc_H, code/calc_h.cpp:9-64 with wpl_part (:66-87) and wab_part
   * (:89-102); BC/AC parts vanish with their fields. */
  cd calc_H() {
    const cd I(0.0, 1.0);
    const double rho0 = p.rho0, kappa = p.kappa;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < M; i++)
      tmp[i] = (kappa <= 0.0 ? cd(0.0) : wpl[i] * wpl[i] * rho0 / kappa / 2.0) -
               I * rho0 * (1.0 - rho_surf[i] - rho_exp_nr[i]) * wpl[i];
    cd h = integ(tmp);
    // Hcur = wpl_part + wbc_part + wac_part + wab_part (:13), pair parts
    // (:89-132) in that order
    auto pair_part = [&](const std::vector<cd>& a, const std::vector<cd>& b, double chi, bool neg) {
      const double sgn = neg ? -1.0 : 1.0;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < M; i++) tmp[i] = (a[i] * a[i] + b[i] * b[i]) * rho0 / chi * sgn;
      return integ(tmp);
    };
    if (chi_bc != 0.0) h += pair_part(wbcp, wbcm, chi_bc, neg_bc);
    if (chi_ac != 0.0) h += pair_part(wacp, wacm, chi_ac, neg_ac);
    if (chi_ab != 0.0) h += pair_part(wabp, wabm, chi_ab, neg_ab);
    if (nD > 0.0) h += -nD * std::log(Qd);
    // chains grafted to the plates (:31-35)
    if (p.n_exp_nr > 0 && p.sigma > 0.0) h += -double(p.n_exp_nr) * ng_per_np * Qga.real();
    Hcur = h;
    return h;
  }
};

/* ------------------------------------------------------------------------ */
/* bcp.input parser, restating read_input (code/read_input.cpp:7-242).       */
/* ------------------------------------------------------------------------ */
namespace {
struct LineReader {
  FILE* f;
  char tt[4096];
  void skip() { if (!fgets(tt, sizeof tt, f)) tt[0] = 0; }
};
}  // namespace

extern "C" int oracle_read_input(const char* path, int dim, scft_params* o) {
  if (dim != 2 && dim != 3) return SCFT_EINVAL;
  FILE* f = fopen(path, "r");
  if (!f) return SCFT_EIO;
  LineReader r{f, {0}};
  memset(o, 0, sizeof *o);
  o->dim = dim;
  int ok = 1;
  auto rd = [&](const char* fmt, void* a) { ok &= fscanf(f, fmt, a) == 1; };
  r.skip();
  rd("%d", &o->Nda); rd("%d", &o->Ndb); r.skip();
  rd("%lf", &o->phiH); rd("%d", &o->Nah); r.skip();
  rd("%lf", &o->a_smear); r.skip();
  rd("%lf", &o->rho0); r.skip();
  r.skip(); r.skip();
  rd("%lf", &o->chi_ab); r.skip();
  rd("%lf", &o->chi_bc); r.skip();
  rd("%lf", &o->chi_ac); r.skip();
  rd("%lf", &o->kappa); r.skip();
  r.skip(); r.skip();
  for (int j = 0; j < 2; j++) {
    rd("%d", &o->ic_flag[j]); rd("%lf", &o->ic_pre[j]);
    rd("%lf", &o->ic_dir[j]); rd("%lf", &o->ic_period[j]); r.skip();
  }
  r.skip(); r.skip();
  for (int i = 0; i < dim; i++) rd("%d", &o->Nx[i]);
  r.skip();
  for (int i = 0; i < dim; i++) rd("%lf", &o->L[i]);
  r.skip();
  rd("%lf", &o->lam_imag); r.skip();
  rd("%lf", &o->lam_real); r.skip();
  rd("%d", &o->itermax); r.skip();
  rd("%d", &o->print_freq); r.skip();
  rd("%d", &o->sample_freq); rd("%d", &o->sample_wait); r.skip();
  rd("%le", &o->error_tol); r.skip();
  rd("%d", &o->update_scheme); r.skip();
  rd("%d", &o->do_CL); r.skip();
  r.skip(); r.skip();
  rd("%d", &o->do_film); r.skip();
  rd("%lf", &o->wallT); rd("%lf", &o->wallXi); r.skip();
  rd("%lf", &o->top_wall_lamA); rd("%lf", &o->top_wall_lamB); r.skip();
  rd("%lf", &o->bot_wall_lamA); rd("%lf", &o->bot_wall_lamB); r.skip();
  r.skip(); r.skip();
  rd("%d", &o->do_channel); rd("%d", &o->hor_dir); rd("%d", &o->vert_dir); r.skip();
  rd("%lf", &o->wall_width); rd("%lf", &o->channel_width); rd("%lf", &o->channel_xi); r.skip();
  r.skip(); r.skip();
  rd("%d", &o->n_exp_nr); r.skip();
  rd("%d", &o->do_fld_np); r.skip();
  rd("%d", &o->np_type); r.skip();
  rd("%d", &o->np_chem); r.skip();
  rd("%lf", &o->sigma); r.skip();
  rd("%d", &o->Ng); r.skip();
  if (o->Ng == 0) o->sigma = 0;
  rd("%d", &o->Nu); r.skip();
  rd("%lf", &o->np_frac); r.skip();
  if (!o->do_fld_np) o->np_frac = 0.0;
  rd("%lf", &o->L_nr); rd("%lf", &o->R_nr); rd("%lf", &o->xi_nr); r.skip();
  for (int j = 0; j < 2; j++) {
    for (int i = 0; i < dim; i++) rd("%lf", &o->exp_nr_c[j][i]);
    r.skip();
    double s = 0;
    for (int i = 0; i < dim; i++) { rd("%lf", &o->exp_nr_u[j][i]); s += o->exp_nr_u[j][i] * o->exp_nr_u[j][i]; }
    r.skip();
    s = std::sqrt(s);
    if (s > 0) for (int i = 0; i < dim; i++) o->exp_nr_u[j][i] /= s;
  }
  fclose(f);
  return ok ? SCFT_OK : SCFT_EINVAL;
}

/* ------------------------------------------------------------------------ */
/* C entry points used by the tests and the CPU-baseline leg of bench.py.    */
/* Arrays are interleaved complex doubles of length M.                        */
/* ------------------------------------------------------------------------ */
extern "C" {

void* oracle_create(const scft_params* p) {
  if (!p || (p->dim != 2 && p->dim != 3) || p->Nda + p->Ndb < 2) return nullptr;
  Oracle* o = new Oracle();
  o->p = *p;
  o->rng.idum = -(long)((p->seed % 2147483000ULL) + 1);
  o->initialize_1();
  o->allocate();
  o->initialize_2();
  return o;
}

void oracle_destroy(void* h) {
  Oracle* o = (Oracle*)h;
  if (!o) return;
  delete o->fft;
  delete o;
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static std::vector<cd>* oracle_array(Oracle* o, const char* name) {
  std::string s(name);
  if (s == "wpl") return &o->wpl;
  if (s == "wabp") return &o->wabp;
  if (s == "wabm") return &o->wabm;
  if (s == "rhoda") return &o->rhoda;
  if (s == "rhodb") return &o->rhodb;
  if (s == "smwa") return &o->smwa;
  if (s == "smwb") return &o->smwb;
  if (s == "rho_exp_nr") return &o->rho_exp_nr;
  if (s == "exp_nrH") return &o->exp_nrH;
  if (s == "hhat") return &o->hhat;
  if (s == "bond") return &o->bond;
  if (s == "gaa") return &o->gaa;
  if (s == "gab") return &o->gab;
  if (s == "gbb") return &o->gbb;
  if (s == "gc") return &o->gc;
  if (s == "wbcp") return &o->wbcp;
  if (s == "wbcm") return &o->wbcm;
  if (s == "wacp") return &o->wacp;
  if (s == "wacm") return &o->wacm;
  if (s == "smwg") return &o->smwg;
  if (s == "expl_grafts") return &o->expl_grafts;
  if (s == "rhog_exp") return &o->rhog_exp;
  if (s.rfind("qg:", 0) == 0) { int j = atoi(name + 3); if (j >= 0 && j < (int)o->qg.size()) return &o->qg[j]; }
  if (s.rfind("q:", 0) == 0) { int j = atoi(name + 2); if (j >= 0 && j <= o->N) return &o->q[j]; }
  if (s.rfind("qdag:", 0) == 0) { int j = atoi(name + 5); if (j >= 0 && j <= o->N) return &o->qdag[j]; }
  return nullptr;
}

int oracle_get_array(void* h, const char* name, double* out) {
  Oracle* o = (Oracle*)h;
  std::vector<cd>* a = oracle_array(o, name);
  if (!a) return SCFT_EINVAL;
  std::memcpy(out, a->data(), sizeof(cd) * o->M);
  return SCFT_OK;
}

int oracle_set_array(void* h, const char* name, const double* in) {
  Oracle* o = (Oracle*)h;
  std::vector<cd>* a = oracle_array(o, name);
  if (!a) return SCFT_EINVAL;
  std::memcpy(a->data(), in, sizeof(cd) * o->M);
  return SCFT_OK;
}

int oracle_set_fields(void* h, const double* wpl, const double* wabp, const double* wabm) {
  Oracle* o = (Oracle*)h;
  if (wpl) std::memcpy(o->wpl.data(), wpl, sizeof(cd) * o->M);
  if (wabp) std::memcpy(o->wabp.data(), wabp, sizeof(cd) * o->M);
  if (wabm) std::memcpy(o->wabm.data(), wabm, sizeof(cd) * o->M);
  return SCFT_OK;
}

double oracle_get_scalar(void* h, const char* name) {
  Oracle* o = (Oracle*)h;
  std::string s(name);
  if (s == "M") return (double)o->M;
  if (s == "N") return o->N;
  if (s == "V") return o->V;
  if (s == "Vf") return o->Vf;
  if (s == "nD") return o->nD;
  if (s == "np_frac") return o->np_frac;
  if (s == "phiD") return o->phiD;
  if (s == "phiG") return o->phiG;
  if (s == "nP") return o->nP;
  if (s == "ng_per_np") return o->ng_per_np;
  if (s == "V_1_exp_np") return o->V_1_exp_np;
  if (s == "Rg2") return o->Rg2;
  if (s == "Qd_re") return o->Qd.real();
  if (s == "Qd_im") return o->Qd.imag();
  if (s == "Qga_re") return o->Qga.real();
  if (s == "Qga_im") return o->Qga.imag();
  return NAN;
}

void oracle_calc_poly_density(void* h) { ((Oracle*)h)->calc_poly_density(); }
void oracle_update_1s(void* h) { ((Oracle*)h)->update_1s(); }
void oracle_update_euler(void* h) { ((Oracle*)h)->update_euler(); }

int oracle_calc_H(void* h, double* out) {
  cd H = ((Oracle*)h)->calc_H();
  out[0] = H.real();
  out[1] = H.imag();
  return std::isnan(H.real()) || std::isnan(H.imag()) ? SCFT_ENAN : SCFT_OK;
}

void oracle_integ(void* h, const double* in, double* out) {
  Oracle* o = (Oracle*)h;
  std::vector<cd> a(o->M);
  std::memcpy(a.data(), in, sizeof(cd) * o->M);
  cd s = o->integ(a);
  out[0] = s.real();
  out[1] = s.imag();
}

/* fft_fwd_wrapper (dir=-1, scaled 1/M) / fft_bck_wrapper (dir=+1). */
void oracle_fft(void* h, const double* in, double* out, int dir) {
  Oracle* o = (Oracle*)h;
  if (dir < 0) o->fft_fwd((const cd*)in, (cd*)out);
  else o->fft_bck((const cd*)in, (cd*)out);
}

/* Unscaled 1D transform of length n (sign -1 forward). */
void oracle_fft1d(int n, const double* in, double* out, int sign) {
  Fft1 f(n);
  std::vector<cd> x(n), s(n);
  std::memcpy(x.data(), in, sizeof(cd) * n);
  f.run(x.data(), s.data(), sign);
  std::memcpy(out, x.data(), sizeof(cd) * n);
}

/* Unnormalised 3D transform on an nx*ny*nz grid (sign -1 forward). */
void oracle_fft3d(int nx, int ny, int nz, double* inout, int sign) {
  Fft3 f(nx, ny, nz);
  f.run((cd*)inout, sign);
}

/* One print row of simulate() (code/simulate.cpp:85-157); the loop itself,
 * with its convergence stop.  rows receives up to max_rows records of
 * {iter, ReH, ImH, Re(-log Qd), Im(-log Qd), err}; returns the row count.
 * If data_path is non-null, data.dat rows are written in the reference's
 * format (code/simulate.cpp:132-135). */
int oracle_simulate(void* h, int max_rows, double* rows, const char* data_path) {
  Oracle* o = (Oracle*)h;
  FILE* otp = data_path ? fopen(data_path, "w") : nullptr;
  o->calc_poly_density();
  o->calc_H();
  cd Hprev_local = 0.0;  // simulate()'s local Hcur starts at zero
  int close = 0, nrows = 0;
  double error = 0.0;
  for (o->iter = 0; o->iter <= o->p.itermax; o->iter++) {
    if (o->p.update_scheme == 0) o->update_euler();
    else o->update_1s();
    if (o->iter % o->p.print_freq == 0) {
      cd Ho = Hprev_local;
      cd H = o->calc_H();
      Hprev_local = H;
      if (H != H) { if (otp) fclose(otp); return SCFT_ENAN; }
      error = std::abs(H - Ho) / o->V / double(o->p.print_freq);
      if (error < o->p.error_tol * 10) close++;
      cd mlq = -std::log(o->Qd);
      if (nrows < max_rows && rows) {
        double* r = rows + 6 * nrows;
        r[0] = o->iter; r[1] = H.real(); r[2] = H.imag();
        r[3] = mlq.real(); r[4] = mlq.imag(); r[5] = error;
      }
      nrows++;
      if (otp) {
        // Qha = Qp = 1 and smwp_min = 0 here (code/calc_poly_density.cpp:122,151)
        cd lqha = -std::log(cd(1.0));
        cd lqp = -std::log(cd(1.0)) + cd(0.0);
        fprintf(otp, "%d %5.6lf %1.3e %5.6lf %1.3e %5.6lf %1.3e %5.6lf %1.3e %1.3e\n",
                o->iter, H.real(), H.imag(), mlq.real(), mlq.imag(), lqha.real(),
                lqha.imag(), lqp.real(), lqp.imag(), error);
      }
    }
    if (!o->p.do_CL && o->iter > 25 && error < o->p.error_tol && close > 10) break;
  }
  if (otp) fclose(otp);
  return nrows;
}

}  // extern "C"
