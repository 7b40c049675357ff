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
// wpl.res, wabp.res, wabm.res, wbcp.res, wbcm.res, wacp.res, wacm.res when
// present (code/io_utils.cpp:138-240: one point per line, Dim coordinates
// then Re Im, x fastest).  Output: the LOG lines of a one-rank run on stdout
// (read_input, initialize_1/2, allocate and simulate reports), data.dat rows
// (code/simulate.cpp:132-135), and <field>.p0.bin dumps every print_freq
// iterations in the reference's binary layout (code/io_utils.cpp:375-397).
// Exit status 1 on a NaN H after writing crashed.* dumps
// (code/simulate.cpp:100-111).
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
  bool grafted = false, pairs = false;
};

int die(const char* what, int rc) {
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, scft_last_error());
  exit(1);
}

// .res reader, read_one_resume_file (code/io_utils.cpp:138-176): every
// point "x [y [z]] re im", blank lines skipped; prints the coordinates of the
// first point as the reference does.
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
    if (i == 0) {
      printf("Processor 0 reading from line 1. x = [ ");
      for (int k = 0; k < dim; k++) printf("%lf ", v[k]);
      printf("]\n");
    }
    out[i] = cd(v[dim], v[dim + 1]);
  }
  fclose(f);
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

// the plate density and graft sites (rods) as initialize_2 writes them
// (code/initialize.cpp:201-238) and write_outputs repeats at iter == 1
void write_setup(const Run& r) {
  std::vector<cd> a(r.M);
  scft_get_cavity(r.ctx, 0, (double*)a.data());
  write_bin(r, "rho_exp_nr", a);
  if (scft_get_array(r.ctx, 0, "expl_grafts", (double*)a.data()) == SCFT_OK) write_bin(r, "expl_grafts", a);
}

// write_outputs, code/io_utils.cpp:11-104.  rhog (field-particle grafts)
// is zero on this path; expl_grafts and rho_exp_nr are written at iter == 1
// (:57-60) besides the set-up copies; the CL averages are the sums divided by
// n_samples (:350-372), which is 0 / 0 before the first sample as there.
void write_outputs(const Run& r, const char* prefix, int iter) {
  std::vector<cd> a(r.M), b(r.M), c(r.M);
  std::string pre(prefix);
  const bool sigma = r.p.sigma > 0.0;
  if (pre.empty()) {
    scft_get_density(r.ctx, 0, (double*)a.data(), (double*)b.data());
    write_bin(r, "rhoda_c", a);
    write_bin(r, "rhodb_c", b);
    smear_host(r, a);
    smear_host(r, b);
    write_bin(r, "rhoda", a);
    write_bin(r, "rhodb", b);
    if (sigma) {
      std::vector<cd> z(r.M, cd(0.0, 0.0));
      write_bin(r, "rhog", z);
      write_bin(r, "rhog_c", z);
      if (r.grafted) {
        scft_get_array(r.ctx, 0, "rhog_exp", (double*)a.data());
      } else {
        a = z;
      }
      write_bin(r, "rhog_exp_c", a);
      smear_host(r, a);
      write_bin(r, "rhog_exp", a);
    }
    if (r.p.n_exp_nr > 0 && iter == 1) write_setup(r);
    if (r.p.do_CL && iter >= r.p.sample_wait) {
      double n = 0.0;
      scft_get_averages(r.ctx, 0, (double*)a.data(), (double*)b.data(), &n);
      // scft_get_averages divides by n once n > 0; before the first sample
      // the reference writes the zero sums divided by n = 0
      if (n == 0.0)
        for (int i = 0; i < r.M; i++) {
          a[i] /= n;
          b[i] /= n;
        }
      write_bin(r, "avg_rhoda", a);
      write_bin(r, "avg_rhodb", b);
      // graft densities are zero on a complex-Langevin path
      const std::vector<cd> z(r.M, cd(0.0, 0.0) / (n == 0.0 ? 0.0 : n));
      write_bin(r, "avg_rhog", z);
      write_bin(r, "avg_rhog_exp", z);
    }
  }
  const bool crash = !pre.empty();
  scft_get_fields(r.ctx, 0, (double*)a.data(), (double*)b.data(), (double*)c.data());
  write_bin(r, (pre + "wpl").c_str(), a);
  if (r.p.chi_ab != 0.0 || crash) {
    write_bin(r, (pre + "wabp").c_str(), b);
    write_bin(r, (pre + "wabm").c_str(), c);
  }
  // the pair fields (zero when the system holds none); crash dumps hold all
  // seven (code/simulate.cpp:103-109)
  const char* pair_names[2][2] = {{"wacp", "wacm"}, {"wbcp", "wbcm"}};
  const double pair_chi[2] = {r.p.chi_ac, r.p.chi_bc};
  for (int k = 0; k < 2; k++) {
    if (pair_chi[k] == 0.0 && !crash) continue;
    for (int j = 0; j < 2; j++) {
      std::fill(a.begin(), a.end(), cd(0.0, 0.0));
      if (r.pairs) scft_get_array(r.ctx, 0, pair_names[k][j], (double*)a.data());
      write_bin(r, (pre + pair_names[k][j]).c_str(), a);
    }
  }
}

double integ_real(const Run& r, const std::vector<cd>& a) {
  cd s = 0.0;
  for (int i = 0; i < r.M; i++) s += a[i];
  for (int d = 0; d < r.p.dim; d++) s *= r.info.dx[d];
  return s.real();
}

// initialize_1's report (code/initialize.cpp:45-81)
void print_init1(const Run& r) {
  printf("\n---Initialization 1---\n\n");
  double V = 1.0;
  for (int d = 0; d < r.p.dim; d++) V *= r.p.L[d];
  const int N = r.p.Nda + r.p.Ndb;
  printf("V = %lf\nN = %d\nfD = %lf\na^2 = %lf\n", V, N, N > 0 ? double(r.p.Nda) / N : 0.0,
         r.p.a_smear * r.p.a_smear);
  printf("---Initialization 1 complete---\n\n");
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
  // read_input (code/read_input.cpp:7-242) and its report
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
  const double chis[3] = {r.p.chi_ab, r.p.chi_bc, r.p.chi_ac};
  const char* chi_names[3] = {"chi_ab", "chi_bc", "chi_ac"};
  for (int k = 0; k < 3; k++)
    if (chis[k] < 0.0) printf("Absolute valuing %s\n", chi_names[k]);
  // SHOW() prints through std::cout: six significant digits (%g)
  for (int k = 0; k < 3; k++) printf("%s = %g\n", chi_names[k], std::fabs(chis[k]));
  for (int k = 0; k < 3; k++) printf("negative_%s_flag = %d\n", chi_names[k], chis[k] < 0.0 ? 1 : 0);
  printf("kappa = %g\n", r.p.kappa);
  for (int i = 0; i < dim; i++) printf("Nx%d: %d Lx%d: %lf\n", i, r.p.Nx[i], i, r.p.L[i]);
  if (r.p.do_fld_np) {
    printf("No field-based nanoparticles!\n");
    return 1;
  }
  printf("---Reading of bcp.input complete---\n");
  print_init1(r);
  rc = scft_create(&r.p, 1, device, prec, &r.ctx);
  if (rc) die("scft_create", rc);
  scft_get_info(r.ctx, 0, &r.info);
  r.M = r.info.M;
  r.grafted = r.p.sigma > 0.0 && r.p.Ng > 0 && r.p.n_exp_nr > 0;
  r.pairs = r.grafted || r.p.chi_bc != 0.0 || r.p.chi_ac != 0.0;
  {
    // allocate() reports its host allocation (code/array_utils.cpp:338-546);
    // the analogue here is the context's device memory
    char buf[8192];
    double mb = 0.0;
    if (scft_describe(r.ctx, buf, sizeof buf) == SCFT_OK) {
      const char* d = strstr(buf, "device_bytes: ");
      if (d) mb = atof(d + 14) / 1.0e6;
    }
    printf("---Allocating memory---\nProcessor 0 allocated %lf MB\n---Memory allocation complete---\n\n", mb);
  }
  printf("------Starting simulation------\n");
  print_init1(r);
  printf("---Initialization 2---\n");
  // set-up copies of the plate density and graft sites (code/initialize.cpp:
  // 201-238)
  if (r.p.n_exp_nr > 0) write_setup(r);
  // .res override of the start fields, read_resume_files
  // (code/io_utils.cpp:183-240); the reference announces the four pair
  // files as "Read wabp.res" / "Read wabm.res", and so does this driver
  {
    const char* files[7] = {"wpl", "wabp", "wabm", "wbcp", "wbcm", "wacp", "wacm"};
    const char* said[7] = {"wpl", "wabp", "wabm", "wabp", "wabm", "wabp", "wabm"};
    std::vector<cd> w;
    for (int k = 0; k < 7; k++) {
      const std::string fn = std::string(files[k]) + ".res";
      if (!read_res(fn.c_str(), dim, r.M, w)) continue;
      if (k >= 3 && !r.pairs) {
        printf("Read %s.res\n", said[k]);
        continue;  // no BC/AC fields in this system: nothing they enter
      }
      rc = scft_set_array(r.ctx, 0, files[k], (const double*)w.data());
      if (rc) die("scft_set_array", rc);
      printf("Read %s.res\n", said[k]);
    }
  }
  printf("Initialized fields\n");
  {
    const scft_sysinfo& f = r.info;
    const double np_frac = f.np_frac;
    printf("Total V_segment actual: %lf\n", f.nD * f.N + f.nP * f.ng_per_np * r.p.Ng);
    printf("Total V_segment theoretical: %lf\n", r.p.rho0 * f.Vf * (1.0 - np_frac));
    printf("V - Vf: %lf\n", f.V - f.Vf);
    printf("rho0 = %lf\n", r.p.rho0);
    printf("V_1_exp_np = %lf\n", f.V_1_exp_np);
    printf("V_1_fld_np = %lf\n", 0.0);
    printf("nD = %lf\n", f.nD);
    printf("nP = %lf\n", f.nP);
    printf("n_exp_nr = %d\n", r.p.n_exp_nr);
    printf("nFP = %lf\n", 0.0);
    printf("Volume fraction of particles excluding grafts = %lf\n", np_frac);
    printf("Volume fraction of grafts                     = %lf\n", f.phiG);
    printf("Volume fraction of particles including grafts = %lf\n", np_frac + f.phiG);
    printf("# of grafts per np = %lf\n", f.ng_per_np);
    printf("Setting up Gaussian bonds\n");
    printf("---Initialization 2 complete---\n\n");
  }
  rc = scft_calc_poly_density(r.ctx);
  if (rc) die("scft_calc_poly_density", rc);
  {
    std::vector<cd> a(r.M), b(r.M);
    scft_get_density(r.ctx, 0, (double*)a.data(), (double*)b.data());
    for (int i = 0; i < r.M; i++) a[i] += b[i];
    printf("Initial densities calculated!\nSegment counts:\n");
    printf("nD * N = %lf integ(rhoda + rhodb) = %lf\n", r.info.nD * r.info.N, integ_real(r, a));
    printf("nAH * Nah = %lf integ(rhoha) = %lf\n", 0.0, 0.0);
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
      if (r.p.sigma > 0.0) {
        double qga[2] = {0.0, 0.0};
        if (r.grafted) scft_get_Qga(r.ctx, 0, qga);
        printf(", -log(Qg)=%lf", -double(r.p.n_exp_nr) * r.info.ng_per_np * qga[0]);
      }
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
