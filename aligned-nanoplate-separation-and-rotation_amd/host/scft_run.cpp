// scft_run: host driver over libscft_amd.so, a drop-in for the reference
// executable's main() + simulate() (code/main.cpp:7-64, code/simulate.cpp:
// 16-167) on one GPU.
//
//   scft_run [--dim 2|3] [--fp32] [--device D] [--seed S] [--no-bin]
//
// The noise seed defaults to the clock, as the reference seeds ran2 from
// time(0) (code/main.cpp:20); --seed makes a complex-Langevin run repeatable.
// Run in the directory holding bcp.input (or bcp.input2), as the reference
// is.  Start fields: init_fields from the input's flags, overridden by
// wpl.res / wabp.res / wabm.res when present (code/io_utils.cpp:138-240:
// one point per line, Dim coordinates then Re Im, x fastest).  Output: LOG
// lines on stdout, data.dat rows (code/simulate.cpp:132-135), and
// <field>.p0.bin dumps every print_freq iterations in the reference's
// binary layout (code/io_utils.cpp:375-397).  Exit status 1 on a NaN H
// after writing crashed.* dumps (code/simulate.cpp:100-111).
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/scft.h"

typedef std::complex<double> cd;

namespace {

struct Run {
  scft_params p{};
  scft_sysinfo info{};
  scft_ctx* ctx = nullptr;
  int M = 0;
  int write_bin = 1;
};

int die(const char* what, int rc) {
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, scft_last_error());
  exit(1);
}

// .res reader: every line "x [y [z]] re im"; blank lines are skipped.
bool read_res(const char* path, int dim, int M, std::vector<cd>& out) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  out.assign(M, cd(0.0, 0.0));
  double v[8];
  for (int i = 0; i < M; i++) {
    for (int k = 0; k < dim + 2; k++) {
      if (fscanf(f, "%lf", &v[k]) != 1) {
        fclose(f);
        fprintf(stderr, "%s: short file (%d of %d points)\n", path, i, M);
        exit(1);
      }
    }
    out[i] = cd(v[dim], v[dim + 1]);
  }
  fclose(f);
  printf("Read %s\n", path);
  return true;
}

// <name>.p0.bin: int Dim, int Nx[Dim], double L[Dim], int nprocs, int ML,
// then ML complex<double> (code/io_utils.cpp:375-397, one rank).
void write_bin(const Run& r, const char* name, const std::vector<cd>& a) {
  if (!r.write_bin) return;
  std::string fn = std::string(name) + ".p0.bin";
  FILE* f = fopen(fn.c_str(), "wb");
  if (!f) return;
  int dim = r.p.dim, nprocs = 1, ml = r.M;
  fwrite(&dim, sizeof(int), 1, f);
  fwrite(r.p.Nx, sizeof(int), dim, f);
  fwrite(r.p.L, sizeof(double), dim, f);
  fwrite(&nprocs, sizeof(int), 1, f);
  fwrite(&ml, sizeof(int), 1, f);
  fwrite(a.data(), sizeof(cd), a.size(), f);
  fclose(f);
}

// smeared density IFFT[h FFT[rho]] for the rhoda / rhodb dumps
// (code/io_utils.cpp:15-31)
void smear_host(const Run& r, std::vector<cd>& a) {
  int rc = scft_fft(r.ctx, 0, (double*)a.data(), -1);
  if (rc) die("scft_fft", rc);
  const int nx = r.p.Nx[0], ny = r.p.Nx[1], nz = r.p.dim == 3 ? r.p.Nx[2] : 1;
  const int N[3] = {nx, ny, nz};
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < r.M; i++) {
    const int n[3] = {i % nx, (i / nx) % ny, i / (nx * ny)};
    double k2 = 0.0;
    for (int d = 0; d < r.p.dim; d++) {
      const double k = (double(n[d]) < double(N[d]) / 2.0) ? 2 * pi * n[d] / r.p.L[d]
                                                           : 2 * pi * (n[d] - N[d]) / r.p.L[d];
      k2 += k * k;
    }
    a[i] *= std::exp(-k2 * r.p.a_smear * r.p.a_smear / 2.0);
  }
  rc = scft_fft(r.ctx, 0, (double*)a.data(), +1);
  if (rc) die("scft_fft", rc);
}

// write_outputs, code/io_utils.cpp:11-104, for the diblock, plates and AB
// fields.  The reference writes rho_exp_nr only when iter == 1 (:57-60) and
// the CL averages divided by n_samples (:350-372); the graft and
// field-particle files have no data on this path.
void write_outputs(const Run& r, const char* prefix, int iter) {
  std::vector<cd> a(r.M), b(r.M), c(r.M);
  std::string pre(prefix);
  if (pre.empty()) {
    scft_get_density(r.ctx, 0, (double*)a.data(), (double*)b.data());
    write_bin(r, "rhoda_c", a);
    write_bin(r, "rhodb_c", b);
    smear_host(r, a);
    smear_host(r, b);
    write_bin(r, "rhoda", a);
    write_bin(r, "rhodb", b);
    if (r.p.n_exp_nr > 0 && iter == 1) {
      scft_get_cavity(r.ctx, 0, (double*)a.data());
      write_bin(r, "rho_exp_nr", a);
    }
    if (r.p.do_CL && iter >= r.p.sample_wait) {
      double n = 0.0;
      scft_get_averages(r.ctx, 0, (double*)a.data(), (double*)b.data(), &n);
      write_bin(r, "avg_rhoda", a);
      write_bin(r, "avg_rhodb", b);
    }
  }
  scft_get_fields(r.ctx, 0, (double*)a.data(), (double*)b.data(), (double*)c.data());
  write_bin(r, (pre + "wpl").c_str(), a);
  // crash dumps hold every field (code/simulate.cpp:103-109)
  if (r.p.chi_ab != 0.0 || !pre.empty()) {
    write_bin(r, (pre + "wabp").c_str(), b);
    write_bin(r, (pre + "wabm").c_str(), c);
  }
}

double integ_real(const Run& r, const std::vector<cd>& a) {
  cd s = 0.0;
  for (int i = 0; i < r.M; i++) s += a[i];
  for (int d = 0; d < r.p.dim; d++) s *= r.info.dx[d];
  return s.real();
}

}  // namespace

int main(int argc, char** argv) {
  Run r;
  int dim = 3, prec = SCFT_PREC_FP64, device = 0, have_seed = 0;
  unsigned long long seed = 0;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--dim") && i + 1 < argc) dim = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--fp32")) prec = SCFT_PREC_FP32;
    else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--seed") && i + 1 < argc) {
      seed = strtoull(argv[++i], nullptr, 10);
      have_seed = 1;
    } else if (!strcmp(argv[i], "--no-bin")) r.write_bin = 0;
    else {
      fprintf(stderr, "usage: scft_run [--dim 2|3] [--fp32] [--device D] [--seed S] [--no-bin]\n");
      return 2;
    }
  }
  printf("---Reading bcp.input---\n");
  int rc = scft_read_input("bcp.input", dim, &r.p);
  if (rc == SCFT_EIO) {
    printf("Failed to open bcp.input!\nTrying bcp2\n");
    rc = scft_read_input("bcp.input2", dim, &r.p);
  }
  if (rc) die("scft_read_input", rc);
  r.p.seed = have_seed ? seed : (unsigned long long)time(nullptr);
  if (r.p.phiH > 0) {
    printf("No homopolymer allowed unless it's a diblock homopolymer!!!!\n");
    return 1;
  }
  if (r.p.do_fld_np) {
    printf("No field-based nanoparticles!\n");
    return 1;
  }
  for (int i = 0; i < dim; i++) printf("Nx%d: %d Lx%d: %lf\n", i, r.p.Nx[i], i, r.p.L[i]);
  rc = scft_create(&r.p, 1, device, prec, &r.ctx);
  if (rc) die("scft_create", rc);
  scft_get_info(r.ctx, 0, &r.info);
  r.M = r.info.M;
  printf("V = %lf\nN = %d\nfD = %lf\na^2 = %lf\n", r.info.V, r.info.N, r.info.fD, r.p.a_smear * r.p.a_smear);
  // .res override of the start fields
  {
    std::vector<cd> wpl, wabp, wabm;
    const bool a = read_res("wpl.res", dim, r.M, wpl);
    const bool b = read_res("wabp.res", dim, r.M, wabp);
    const bool c = read_res("wabm.res", dim, r.M, wabm);
    if (a || b || c) {
      rc = scft_set_fields(r.ctx, 0, a ? (double*)wpl.data() : nullptr, b ? (double*)wabp.data() : nullptr,
                           c ? (double*)wabm.data() : nullptr);
      if (rc) die("scft_set_fields", rc);
    }
  }
  printf("Total V_segment actual: %lf\n", r.info.nD * r.info.N);
  printf("V_1_exp_np = %lf\nnD = %lf\n", r.info.V_1_exp_np, r.info.nD);
  printf("Volume fraction of particles excluding grafts = %lf\n", r.info.np_frac);
  rc = scft_calc_poly_density(r.ctx);
  if (rc) die("scft_calc_poly_density", rc);
  {
    std::vector<cd> a(r.M), b(r.M);
    scft_get_density(r.ctx, 0, (double*)a.data(), (double*)b.data());
    for (int i = 0; i < r.M; i++) a[i] += b[i];
    printf("Initial densities calculated!\nSegment counts:\n");
    printf("nD * N = %lf integ(rhoda + rhodb) = %lf\n", r.info.nD * r.info.N, integ_real(r, a));
  }
  double h[2];
  rc = scft_calc_H(r.ctx, h);
  if (rc && rc != SCFT_ENAN) die("scft_calc_H", rc);
  printf("Starting H: %lf\n\n", h[0]);
  write_outputs(r, "", 0);
  printf("---Entering main loop---\n");
  fflush(stdout);

  FILE* otp = fopen("data.dat", "w");
  if (!otp) {
    printf("Failed to open data.dat!\n");
    return 1;
  }
  cd Hloc = 0.0;  // simulate()'s local Hcur starts at zero
  double error = 0.0;
  int close = 0;
  const int pf = r.p.print_freq > 0 ? r.p.print_freq : 1;
  int iter = 0;
  while (iter <= r.p.itermax) {
    // iterations up to the next print (or the end) in one graph-replay call
    int next = ((iter / pf) + (iter % pf ? 1 : 0)) * pf;
    if (next > r.p.itermax) next = r.p.itermax;
    const int n = next - iter + 1;
    rc = scft_iterate(r.ctx, n);
    if (rc) die("scft_iterate", rc);
    iter = next;
    if (iter % pf == 0) {
      const cd Ho = Hloc;
      rc = scft_calc_H(r.ctx, h);
      const cd H(h[0], h[1]);
      Hloc = H;
      double Q[2];
      scft_get_Q(r.ctx, 0, Q);
      if (rc == SCFT_ENAN || H != H) {
        printf("Crashed! iteration %d\n", iter);
        printf("H: %lf Qd: %lf\n", h[0], Q[0]);
        write_outputs(r, "crashed.", iter);
        fclose(otp);
        return 1;
      }
      error = std::abs(H - Ho) / r.info.V / double(pf);
      if (error < r.p.error_tol * 10) close++;
      const cd mlq = -std::log(cd(Q[0], Q[1]));
      printf("Iter: %d, H=%lf", iter, H.real());
      if (r.p.do_CL) printf(" + i%lf", H.imag());
      printf(", -log(Qd)=%lf", mlq.real());
      printf(", err=%1.1e\n", error);
      fflush(stdout);
      // no homopolymer and no field particles: Qha = Qp = 1 and smwp_min = 0
      // (code/calc_poly_density.cpp:122,151), so these columns print -0 and +0
      const cd lqha = -std::log(cd(1.0, 0.0));
      const cd lqp = -std::log(cd(1.0, 0.0)) + cd(0.0, 0.0);
      fprintf(otp, "%d %5.6lf %1.3e %5.6lf %1.3e %5.6lf %1.3e %5.6lf %1.3e %1.3e\n", iter, H.real(), H.imag(),
              mlq.real(), mlq.imag(), lqha.real(), lqha.imag(), lqp.real(), lqp.imag(), error);
      fflush(otp);
      write_outputs(r, "", iter);
    }
    if (!r.p.do_CL && iter > 25 && error < r.p.error_tol && close > 10) {
      printf("Tolerance reached. Error = %.4e\n---Main loop complete---\n\n", error);
      break;
    }
    iter++;
  }
  fclose(otp);
  printf("------Completed simulation------\n\n");
  scft_destroy(r.ctx);
  const time_t now = time(nullptr);
  printf("The program finished at %s.", ctime(&now));
  return 0;
}
