// Host side of libscft_amd.so: context, set-up and the C ABI of include/scft.h.
//
// One context holds n systems (possibly of different box and grid) on one
// GPU and one HIP stream.  Everything the hot path touches lives in HBM:
//   per system, complex128: w+, wAB+, wAB-, rhoA, rhoB, wa/smwa, wb/smwb,
//     rho_np, exp_nrH, an FFT scratch; real128: h(k), X(k), X2(k);
//   per system, propagator precision: q and q-dagger as N*M contiguous
//     arrays (segment-major), exp(-wA), exp(-wB).
// FFT-style work is described once, at create, as "programs": lists of
// launches over device tables of PassDesc.  The contour step is a kernel
// argument, so a whole iteration is a fixed launch sequence that is
// captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/scft.h"
#include "kernels.h"
#include "scft_dev.hpp"

using scft::PassDesc;
using scft::Plan;
using scft::SysDev;
typedef std::complex<double> cd;

namespace {

const double kPi = 3.14159265358979323846264338327950288;
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCK(x)                                                                   \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(SCFT_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));  \
  } while (0)

// ---------------------------------------------------------------------------
// radix plans and twiddles
// ---------------------------------------------------------------------------
bool make_plan(int n, Plan* p) {
  p->n = n;
  p->npass = 0;
  p->radices = 0;
  int m = n;
  const int order[] = {8, 4, 2, 3, 5, 7};
  for (int r : order) {
    while (m % r == 0) {
      if (p->npass == 8) return false;
      p->radices |= (unsigned)r << (4 * p->npass);
      p->npass++;
      m /= r;
    }
  }
  return m == 1;
}

// largest register footprint of a pass: ceil((n B / R) / T) * R values
int pass_values(const Plan& p, int nB, int T) {
  int worst = 0;
  for (int i = 0; i < p.npass; i++) {
    const int R = p.radix(i);
    const int total = nB / R;
    const int u = (total + T - 1) / T;
    worst = std::max(worst, u * R);
  }
  return worst;
}

constexpr int kMaxV = 32;
constexpr size_t kMaxLds = 160 * 1024;

int pick_tb(const Plan& a, const Plan* b, int nB) {
  for (int T : {256, 512, 1024}) {
    if (pass_values(a, nB, T) <= kMaxV && (!b || pass_values(*b, nB, T) <= kMaxV)) return T;
  }
  return 0;
}

// k of index n on an axis of length Nx and box L, code/array_utils.cpp:248-251
inline double kval(int n, int Nx, double L) {
  return (double(n) < double(Nx) / 2.0) ? 2 * kPi * double(n) / L : 2 * kPi * double(n - Nx) / L;
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct Launch {
  bool plane;
  int fast;  // 0 generic, 1 size-specialised plane conv, 2 size-specialised z conv
  int n0, n1;
  int mode, prec;
  PassDesc* d;
  int njobs, gridx, tb;
  size_t lds;
  // size-specialised launches: candidate variants, z tile widths, choice
  std::vector<int> variants, tiles;
  std::vector<float> tuned_ms;
  int choice = 0;
  int max_lines = 0;
};

struct Program {
  std::vector<Launch> launches;
};

struct SysHost {
  scft_params p;
  scft_sysinfo info;
  int nx, ny, nz, M, N, Na;
  double L[3], dx[3], dV;
  int neg_ab;
  double chi_ab;
  SysDev dev;  // host copy of the device table row
  // 1D tables, per axis: bond (prop precision) and smear (FP64), device
  void* bond_g[3];
  void* smear_g[3];
  double2* grad_out = nullptr;  // device, 8 complex
  // BC / AC pairs and grafts
  bool pairs = false, grafted = false, has_sites = false;
  int neg_bc = 0, neg_ac = 0;
  double chi_bc = 0.0, chi_ac = 0.0;
  cd expl_total = 0.0;  // integ(expl_grafts) after normalisation (unit)
  cd expl_Z = 0.0;      // integ(sum_p |grad shape_p|), the normalisation
  std::vector<void*> allocs;
};

struct scft_ctx {
  int device = 0;
  int prec = 0;
  size_t esz = 16;  // bytes per propagator complex
  hipStream_t st = nullptr;
  // second stream of the q-dagger chain: the two chains of the propagator
  // run as parallel branches (fork / join events), so one chain's launches
  // fill the other's gaps and tails
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // third branch: the graft chains, concurrent with the diblock chains
  hipStream_t st3 = nullptr;
  hipEvent_t ev_gjoin = nullptr;
  bool split_chains = false;
  std::vector<SysHost> sys;
  SysDev* d_sys = nullptr;
  int maxM = 0, maxN = 0, nblk = 1;
  std::map<int, size_t> tw_off;  // n -> element offset in the twiddle tables
  void* tw64 = nullptr;
  void* tw32 = nullptr;
  std::vector<void*> allocs;
  Program prog_smear, prog_prop, prog_fwd, prog_inv, prog_noise;
  Program prog_prop_q, prog_prop_qd;  // the two chains of prog_prop, one each
  Program prog_graft_q, prog_graft_qd;  // graft chains (free end, grafted end)
  bool any_graft = false, any_pairs = false;
  size_t dev_bytes = 0;  // device memory held by the context
  int maxNg = 0;
  // checkpointed propagators (SCFT_OPT_CHECKPOINT): segment e covers the q
  // indices ck_lo[e] .. ck_lo[e] + ck_n[e] - 1 counted from the top
  // (e = 0 ends at N - 1); prog_ck_q / prog_ck_qd step the two windows
  bool ckpt = false;
  int ck_K = 0, ck_ne = 0;
  std::vector<int> ck_lo, ck_n;
  std::vector<hipEvent_t> ck_fork, ck_join;
  Program prog_ck_q, prog_ck_qd;
  std::vector<Program> prog_fft_fwd, prog_fft_inv;  // per system, on its scratch
  std::vector<void*> scratch;                       // per system FP64 scratch
  hipGraphExec_t graph = nullptr;
  int graph_scheme = -1;
  bool any_cl = false;
};

namespace {

int dalloc(scft_ctx* c, std::vector<void*>& list, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(SCFT_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  list.push_back(*p);
  c->dev_bytes += bytes;
  return SCFT_OK;
}

// ---------------------------------------------------------------------------
// program building
// ---------------------------------------------------------------------------
struct JobSpec {
  int sys;
  int prec;
  void* in;
  void* out;
  long in_step, out_step;
  int in_soff, out_soff;
  const void* post_a;
  const void* post_b;
  int post_switch;
  int steps_active;
  int gkind;  // 0 none, 1 bond, 2 smear
  double scale;
};

enum StageKind { ST_PLANE = 0, ST_X = 1, ST_Y = 2, ST_Z = 3 };

const void* tw_ptr(scft_ctx* c, int n, int prec) {
  const size_t off = c->tw_off.at(n);
  return prec == 0 ? (const void*)((double2*)c->tw64 + off) : (const void*)((float2*)c->tw32 + off);
}

int build_program(scft_ctx* c, int mode, const std::vector<JobSpec>& jobs, Program* out) {
  // stage lists per job
  struct Item {
    int job;
    int stage;  // position in the job's sequence
    int kind;
    bool last;
  };
  std::vector<std::vector<int>> seq(jobs.size());
  for (size_t j = 0; j < jobs.size(); j++) {
    const SysHost& s = c->sys[jobs[j].sys];
    const size_t esz = jobs[j].prec == 0 ? 16 : 8;
    Plan px, py;
    make_plan(s.nx, &px);
    make_plan(s.ny, &py);
    const bool fits = (size_t)s.nx * s.ny * esz <= kMaxLds &&
                      pick_tb(px, &py, s.nx * s.ny) != 0 && s.nx * s.ny > 1;
    std::vector<int>& q = seq[j];
    if (fits) q.push_back(ST_PLANE);
    else {
      if (s.nx > 1) q.push_back(ST_X);
      if (s.ny > 1) q.push_back(ST_Y);
    }
    if (s.nz > 1) q.push_back(ST_Z);
  }
  // group by (stage index, kind, prec, specialised size)
  std::map<std::tuple<int, int, int, int, int, int>, std::vector<Item>> groups;
  for (size_t j = 0; j < jobs.size(); j++) {
    const SysHost& s = c->sys[jobs[j].sys];
    for (size_t k = 0; k < seq[j].size(); k++) {
      const int kind = seq[j][k];
      int fast = 0, n0 = 0, n1 = 0;
      int ids[16], tls[16];
      if (kind == ST_PLANE && scft::fast_plane_variants(jobs[j].prec, s.nx, s.ny, ids, 16) > 0) {
        fast = 1;
        n0 = s.nx;
        n1 = s.ny;
      } else if (kind == ST_Z && scft::fast_z_variants(jobs[j].prec, s.nz, ids, tls, 16) > 0) {
        fast = 2;
        n0 = s.nz;
      } else if ((kind == ST_X || kind == ST_Y) &&
                 scft::odd_line_variants(jobs[j].prec, kind == ST_X ? s.nx : s.ny, kind - ST_X, ids, tls, 16) > 0) {
        fast = 3;
        n0 = kind == ST_X ? s.nx : s.ny;
        n1 = kind - ST_X;
      }
      groups[{(int)k, kind, jobs[j].prec, fast, n0, n1}].push_back(
          {(int)j, (int)k, kind, k + 1 == seq[j].size()});
    }
  }
  std::vector<PassDesc> all;
  std::vector<Launch> launches;
  std::vector<size_t> starts;
  for (auto& kv : groups) {
    const int kind = std::get<1>(kv.first);
    const int prec = std::get<2>(kv.first);
    const size_t esz = prec == 0 ? 16 : 8;
    Launch L{};
    L.plane = kind == ST_PLANE;
    L.fast = std::get<3>(kv.first);
    L.n0 = std::get<4>(kv.first);
    L.n1 = std::get<5>(kv.first);
    if (L.fast) {
      int ids[16], tls[16];
      const int nv = L.fast == 1   ? scft::fast_plane_variants(prec, L.n0, L.n1, ids, 16)
                     : L.fast == 2 ? scft::fast_z_variants(prec, L.n0, ids, tls, 16)
                                   : scft::odd_line_variants(prec, L.n0, L.n1, ids, tls, 16);
      for (int v = 0; v < nv; v++) {
        L.variants.push_back(ids[v]);
        L.tiles.push_back(L.fast >= 2 ? tls[v] : 0);
      }
    }
    L.mode = mode;
    L.prec = prec;
    L.njobs = (int)kv.second.size();
    starts.push_back(all.size());
    for (const Item& it : kv.second) {
      const JobSpec& J = jobs[it.job];
      const SysHost& s = c->sys[J.sys];
      PassDesc d{};
      if (it.stage == 0) {
        d.in = J.in;
        d.in_step = J.in_step;
        d.in_soff = J.in_soff;
      } else {
        d.in = J.out;
        d.in_step = J.out_step;
        d.in_soff = J.out_soff;
      }
      d.out = J.out;
      d.out_step = J.out_step;
      d.out_soff = J.out_soff;
      d.post_a = J.post_a;
      d.post_b = J.post_b;
      d.post_switch = J.post_switch;
      d.steps_active = J.steps_active;
      d.nx = s.nx;
      d.ny = s.ny;
      d.nz = s.nz;
      d.last = it.last ? 1 : 0;
      d.scale = J.scale;
      auto gsel = [&](int axis) -> const void* {
        if (mode != scft::MODE_CONV) return nullptr;
        return J.gkind == 1 ? s.bond_g[axis] : s.smear_g[axis];
      };
      if (kind == ST_PLANE) {
        make_plan(s.nx, &d.p0);
        make_plan(s.ny, &d.p1);
        d.tw0 = tw_ptr(c, s.nx, prec);
        d.tw1 = tw_ptr(c, s.ny, prec);
        d.g0 = gsel(0);
        d.g1 = gsel(1);
        d.ntiles = s.nz;
        const int tb = pick_tb(d.p0, &d.p1, s.nx * s.ny);
        L.tb = std::max(L.tb, tb);
        L.lds = std::max(L.lds, (size_t)s.nx * s.ny * esz);
        L.gridx = std::max(L.gridx, s.nz);
      } else {
        const int axis = kind - ST_X;
        const int n = axis == 0 ? s.nx : (axis == 1 ? s.ny : s.nz);
        d.axis = axis;
        make_plan(n, &d.p0);
        d.tw0 = tw_ptr(c, n, prec);
        d.g0 = gsel(axis);
        const int nlines = s.M / n;
        if (L.fast >= 2) {
          L.max_lines = std::max(L.max_lines, nlines);
          all.push_back(d);
          continue;
        }
        const int budget = prec == 0 ? 2048 : 4096;
        int tl = std::max(1, budget / n);
        if (tl >= 8) tl &= ~7;
        tl = std::min(tl, 64);
        tl = std::min(tl, nlines);
        int tb = 0;
        while (tl >= 1) {
          tb = pick_tb(d.p0, nullptr, tl * n);
          if (tb && (size_t)tl * n * esz <= kMaxLds) break;
          tl = tl > 8 ? tl - 8 : tl - 1;
        }
        if (tl < 1 || !tb) return fail(SCFT_EUNSUPPORTED, "no tile shape for line length " + std::to_string(n));
        d.tl = tl;
        d.ntiles = (nlines + tl - 1) / tl;
        L.tb = std::max(L.tb, tb);
        L.lds = std::max(L.lds, (size_t)tl * n * esz);
        L.gridx = std::max(L.gridx, d.ntiles);
      }
      all.push_back(d);
    }
    launches.push_back(L);
  }
  // upload the descriptor table
  PassDesc* dd = nullptr;
  if (!all.empty()) {
    void* p;
    int rc = dalloc(c, c->allocs, &p, all.size() * sizeof(PassDesc));
    if (rc) return rc;
    dd = (PassDesc*)p;
    HIPCK(hipMemcpy(dd, all.data(), all.size() * sizeof(PassDesc), hipMemcpyHostToDevice));
  }
  // order launches by stage index (map order is stage-major already)
  for (size_t i = 0; i < launches.size(); i++) launches[i].d = dd + starts[i];
  out->launches = launches;
  return SCFT_OK;
}

// `step` is the slice index of the arrays; the contour index is
// step + contour_off (they differ only in the checkpointed propagator)
hipError_t launch_one(scft_ctx* c, const Launch& L, int step, hipStream_t st = nullptr, int contour_off = 0) {
  if (!st) st = c->st;
  step = scft::pack_step(step, step + contour_off);
  if (L.fast == 1)
    return scft::launch_fast_plane(L.prec, L.mode, L.variants[L.choice], L.d, L.njobs, L.gridx, step, st);
  if (L.fast == 2) {
    const int tl = L.tiles[L.choice];
    return scft::launch_fast_z(L.prec, L.mode, L.variants[L.choice], L.d, L.njobs, (L.max_lines + tl - 1) / tl,
                               step, st);
  }
  if (L.fast == 3) {
    const int cpb = L.tiles[L.choice];
    int gridx = (L.max_lines + cpb - 1) / cpb;
    // strided passes group columns by XCD from the block index; workgroups
    // go to XCDs by linear id x + gridx y, so gridx is padded to a multiple
    // of 8 (the extra blocks exit) to keep that mapping for every job y
    if (L.n1 == 1) gridx = (gridx + 7) & ~7;
    return scft::launch_odd_line(L.prec, L.mode, L.variants[L.choice], L.d, L.njobs, gridx, step, st);
  }
  return scft::launch_pass(L.prec, L.mode, L.plane, L.d, L.njobs, L.gridx, L.tb, L.lds, step, st);
}

// Autotuner: time every variant of each size-specialised launch of a
// program on the context's own buffers (step 1 of the propagator recomputes
// q_1 from q_0; the smearing works in place on scratch fields) and keep the
// fastest.
int autotune(scft_ctx* c, Program& p, int step) {
  hipEvent_t e0, e1;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  // SCFT_TUNE_PICK=k: take variant k (or the last) of every launch instead
  // of timing them, so that tests can run each variant
  const char* pick = getenv("SCFT_TUNE_PICK");
  for (Launch& L : p.launches) {
    if (L.variants.size() < 2) continue;
    if (pick) {
      L.choice = std::min(std::max(atoi(pick), 0), (int)L.variants.size() - 1);
      continue;
    }
    float best = 1e30f;
    int best_i = 0;
    for (size_t v = 0; v < L.variants.size(); v++) {
      L.choice = (int)v;
      for (int r = 0; r < 2; r++) HIPCK(launch_one(c, L, step));
      HIPCK(hipEventRecord(e0, c->st));
      for (int r = 0; r < 8; r++) HIPCK(launch_one(c, L, step));
      HIPCK(hipEventRecord(e1, c->st));
      HIPCK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCK(hipEventElapsedTime(&ms, e0, e1));
      L.tuned_ms.push_back(ms / 8.f);
      if (ms < best) {
        best = ms;
        best_i = (int)v;
      }
    }
    L.choice = best_i;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return SCFT_OK;
}

int run_program(scft_ctx* c, const Program& p, int step, hipStream_t st = nullptr, int contour_off = 0) {
  for (const Launch& L : p.launches) HIPCK(launch_one(c, L, step, st, contour_off));
  return SCFT_OK;
}

// ---------------------------------------------------------------------------
// set-up of one system: initialize_1 / initialize_2 (code/initialize.cpp)
// ---------------------------------------------------------------------------
void plate_axes(const double* u, double* e1, double* e2) {
  double c[3] = {0.0, u[2], -u[1]};  // u x X
  double n = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  if (n < 1e-12) {
    c[0] = -u[2];  // u x Y
    c[1] = 0.0;
    c[2] = u[0];
    n = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  }
  for (int d = 0; d < 3; d++) e2[d] = c[d] / n;
  e1[0] = e2[1] * u[2] - e2[2] * u[1];
  e1[1] = e2[2] * u[0] - e2[0] * u[2];
  e1[2] = e2[0] * u[1] - e2[1] * u[0];
}

// Plate density at grid point (n0, n1, n2): explicit_nanorod
// (code/initialize.cpp:578-612) for np_type 2, the square plate otherwise.
double plate_rho(const SysHost& s, int np, const int* nn, double scale = 1.0) {
  const scft_params& p = s.p;
  const double len = scale * p.L_nr, rad = scale * p.R_nr;
  const int dim = p.dim;
  double x[3], c[3], dr[3] = {0, 0, 0};
  for (int d = 0; d < dim; d++) {
    x[d] = double(nn[d]) * s.dx[d];
    c[d] = p.exp_nr_c[np][d] * s.L[d];
    dr[d] = x[d] - c[d];
    if (dr[d] >= 0.5 * s.L[d]) dr[d] -= s.L[d];
    else if (dr[d] < -0.5 * s.L[d]) dr[d] += s.L[d];
  }
  const double* u = p.exp_nr_u[np];
  double ud = 0;
  for (int d = 0; d < dim; d++) ud += u[d] * dr[d];
  const double a = std::erfc((std::fabs(ud) - 0.5 * len) / p.xi_nr);
  if (p.np_type == SCFT_NP_SQUARE && dim == 3) {
    double e1[3], e2[3];
    plate_axes(u, e1, e2);
    const double a1 = e1[0] * dr[0] + e1[1] * dr[1] + e1[2] * dr[2];
    const double a2 = e2[0] * dr[0] + e2[1] * dr[1] + e2[2] * dr[2];
    return 0.125 * a * std::erfc((std::fabs(a1) - rad) / p.xi_nr) *
           std::erfc((std::fabs(a2) - rad) / p.xi_nr);
  }
  double uc;
  if (dim == 2) uc = std::fabs(u[0] * dr[1] - u[1] * dr[0]);
  else {
    const double c0 = u[1] * dr[2] - u[2] * dr[1], c1 = u[2] * dr[0] - u[0] * dr[2],
                 c2 = u[0] * dr[1] - u[1] * dr[0];
    uc = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }
  return 0.25 * a * std::erfc((std::fabs(uc) - rad) / p.xi_nr);
}

// Numerical Recipes ran2 + polar Box-Muller, for ic_flag == -1 starts
// (code/ran2.cpp:32-90 uses the same published generator).
struct Ran2 {
  long idum, idum2 = 123456789, iy = 0, iv[32];
  explicit Ran2(long seed) : idum(seed) {}
  double next() {
    const long m1 = 2147483563, m2 = 2147483399, a1 = 40014, a2 = 40692, q1 = 53668,
               q2 = 52774, r1 = 12211, r2 = 3791, mm1 = m1 - 1, ndiv = 1 + mm1 / 32;
    if (idum <= 0) {
      idum = (-idum < 1) ? 1 : -idum;
      idum2 = idum;
      for (int j = 39; j >= 0; j--) {
        long h = idum / q1;
        idum = a1 * (idum - h * q1) - h * r1;
        if (idum < 0) idum += m1;
        if (j < 32) iv[j] = idum;
      }
      iy = iv[0];
    }
    long h = idum / q1;
    idum = a1 * (idum - h * q1) - h * r1;
    if (idum < 0) idum += m1;
    h = idum2 / q2;
    idum2 = a2 * (idum2 - h * q2) - h * r2;
    if (idum2 < 0) idum2 += m2;
    const int j = (int)(iy / ndiv);
    iy = iv[j] - idum2;
    iv[j] = idum;
    if (iy < 1) iy += mm1;
    const double t = iy * (1.0 / m1);
    return t > 1.0 - 1.2e-7 ? 1.0 - 1.2e-7 : t;
  }
};

// Device row of plate np from the host parameters (absolute centre, axes).
void set_plate_row(SysHost& s, int np) {
  scft::PlateDev& P = s.dev.plate[np];
  const scft_params& p = s.p;
  for (int d = 0; d < 3; d++) {
    P.c[d] = d < p.dim ? p.exp_nr_c[np][d] * s.L[d] : 0.0;
    P.u[d] = d < p.dim ? p.exp_nr_u[np][d] : 0.0;
  }
  plate_axes(P.u, P.e1, P.e2);
}

int setup_system(scft_ctx* c, int si) {
  SysHost& s = c->sys[si];
  const scft_params& p = s.p;
  if (p.dim != 2 && p.dim != 3) return fail(SCFT_EINVAL, "dim must be 2 or 3");
  if (p.Nda < 0 || p.Ndb < 0 || p.Nda + p.Ndb < 2) return fail(SCFT_EINVAL, "Nda + Ndb must be >= 2");
  if (p.do_film || p.do_channel || p.do_fld_np || p.phiH > 0 || p.np_type == SCFT_NP_SPHERE)
    return fail(SCFT_EUNSUPPORTED, "films, channels, field nanoparticles, homopolymer and spheres are not on this path");
  s.grafted = p.sigma > 0.0 && p.Ng > 0 && p.n_exp_nr > 0;
  s.pairs = s.grafted || p.chi_bc != 0.0 || p.chi_ac != 0.0;
  s.has_sites = p.np_type == SCFT_NP_ROD && p.n_exp_nr > 0;
  if (s.grafted && p.np_type != SCFT_NP_ROD)
    return fail(SCFT_EUNSUPPORTED, "grafted chains are built for rod plates (np_type 2) only");
  if (s.pairs && (p.update_scheme != 1 || p.do_CL))
    return fail(SCFT_EUNSUPPORTED, "grafts and BC/AC pairs are built for the 1S update without noise");
  s.neg_bc = p.chi_bc < 0;
  s.chi_bc = std::fabs(p.chi_bc);
  s.neg_ac = p.chi_ac < 0;
  s.chi_ac = std::fabs(p.chi_ac);
  if (p.do_CL && p.sample_freq < 1) return fail(SCFT_EINVAL, "do_CL needs sample_freq >= 1");
  s.nx = p.Nx[0];
  s.ny = p.Nx[1];
  s.nz = p.dim == 3 ? p.Nx[2] : 1;
  if (s.nx < 1 || s.ny < 1 || s.nz < 1) return fail(SCFT_EINVAL, "bad grid");
  s.M = s.nx * s.ny * s.nz;
  s.N = p.Nda + p.Ndb;
  s.Na = p.Nda;
  double V = 1.0;
  s.dV = 1.0;
  for (int d = 0; d < 3; d++) {
    const int n = d == 0 ? s.nx : (d == 1 ? s.ny : s.nz);
    s.L[d] = d < p.dim ? p.L[d] : 1.0;
    s.dx[d] = s.L[d] / double(n);
    if (d < p.dim) {
      V *= s.L[d];
      s.dV *= s.dx[d];
    }
  }
  for (int n : {s.nx, s.ny, s.nz}) {
    Plan pl;
    if (!make_plan(n, &pl)) return fail(SCFT_EUNSUPPORTED, "grid size " + std::to_string(n) + " has a prime factor above 7");
  }
  s.neg_ab = p.chi_ab < 0;
  s.chi_ab = std::fabs(p.chi_ab);
  const int M = s.M, N = s.N;
  const double fD = double(p.Nda) / double(N);
  const double Rg2 = double(N - 1) / 6.0;
  const double a2 = p.a_smear * p.a_smear;

  // plate density and volume bookkeeping, code/initialize.cpp:210-355
  std::vector<cd> rho_np(M, cd(0.0, 0.0));
  for (int np = 0; np < p.n_exp_nr && np < 2; np++) {
    for (int i = 0; i < M; i++) {
      int nn[3] = {i % s.nx, (i / s.nx) % s.ny, i / (s.nx * s.ny)};
      rho_np[i] += plate_rho(s, np, nn);
    }
  }
  cd sum = 0.0;
  for (int i = 0; i < M; i++) sum += rho_np[i];
  for (int d = 0; d < p.dim; d++) sum *= s.dx[d];
  const double V_exp = sum.real();
  const double Vf = V;
  const double np_frac = V_exp / Vf;
  const double V1 = p.n_exp_nr > 0 ? V_exp / double(p.n_exp_nr) : 0.0;
  const double nP = p.n_exp_nr > 0 ? V_exp / V1 : 0.0;
  // grafts, code/initialize.cpp:331-345
  double ng_per_np = 0.0, phiG = 0.0;
  if (p.sigma > 0.0) {
    ng_per_np = p.dim == 2 ? p.sigma * 2.0 * kPi * p.R_nr : p.sigma * 4.0 * kPi * p.R_nr * p.R_nr;
    phiG = ng_per_np * nP * p.Ng / p.rho0 / Vf;
  }
  const double phiD = 1.0 - p.phiH - np_frac - phiG;
  const double nD = phiD * p.rho0 * Vf / N;
  scft_sysinfo& inf = s.info;
  memset(&inf, 0, sizeof inf);
  inf.M = M;
  inf.N = N;
  inf.V = V;
  inf.Vf = Vf;
  for (int d = 0; d < 3; d++) inf.dx[d] = s.dx[d];
  inf.V_1_exp_np = V1;
  inf.np_frac = np_frac;
  inf.phiD = phiD;
  inf.phiG = phiG;
  inf.nP = nP;
  inf.ng_per_np = ng_per_np;
  inf.nD = nD;
  inf.Rg2 = Rg2;
  inf.fD = fD;

  // k-space coefficients: h(k) (code/initialize.cpp:97-102), Debye sums
  // (code/calc_debye.cpp:5-51) folded into X(k) = phiD rho0 N h^2 (gaa + 2 gab + gbb)
  std::vector<double> hk(M), X(M), X2(M), XA(M), XB(M), XG(M);
  const double galpha = p.Ng > 0 ? double(p.Ng - 1) / double(N - 1) : 0.0;
  for (int i = 0; i < M; i++) {
    const int n0 = i % s.nx, n1 = (i / s.nx) % s.ny, n2 = i / (s.nx * s.ny);
    double k2 = 0.0;
    const int nn[3] = {n0, n1, n2};
    const int Ns[3] = {s.nx, s.ny, s.nz};
    for (int d = 0; d < p.dim; d++) {
      const double k = kval(nn[d], Ns[d], s.L[d]);
      k2 += k * k;
    }
    const double h = std::exp(-k2 * a2 / 2.0);
    const double x = k2 * Rg2;
    double gaa, gbb, gab;
    if (i == 0) {
      gaa = fD * fD;
      gbb = (1.0 - fD) * (1.0 - fD);
      gab = fD * (1.0 - fD);
    } else {
      gaa = 2.0 * (std::exp(-x * fD) + fD * x - 1.0) / x / x;
      gbb = 2.0 * (std::exp(-x * (1.0 - fD)) + (1.0 - fD) * x - 1.0) / x / x;
      gab = (1.0 - std::exp(-x * fD)) * (1.0 - std::exp(-x * (1.0 - fD))) / x / x;
    }
    hk[i] = h;
    X[i] = phiD * p.rho0 * N * h * h * (gaa + 2.0 * gab + gbb);
    X2[i] = phiD * p.rho0 * N * h * h * (gaa - gbb);
    XA[i] = phiD * p.rho0 * N * h * h * gaa;
    XB[i] = phiD * p.rho0 * N * h * h * gbb;
    // graft Debye function, calc_gd (code/calc_debye.cpp:53-68) with
    // alpha = (Ng - 1) / (N - 1), code/initialize.cpp:394
    const double dg = x * galpha;
    const double gc = dg == 0.0 ? 1.0 : 2.0 * (std::exp(-dg) + dg - 1.0) / dg / dg;
    XG[i] = phiG * p.rho0 * p.Ng * h * h * gc;
  }

  // initial fields, init_fields (code/initialize.cpp:419-565)
  std::vector<cd> wpl(M, 0.0), wabp(M, 0.0), wabm(M, 0.0);
  std::vector<cd> wbcp, wbcm, wacp, wacm;
  if (s.pairs) {
    wbcp.assign(M, 0.0);
    wbcm.assign(M, 0.0);
    wacp.assign(M, 0.0);
    wacm.assign(M, 0.0);
  }
  {
    const cd I(0.0, 1.0);
    Ran2 rng(-(long)((p.seed % 2147483000ULL) + 1));
    for (int i = 0; i < M; i++) {
      const int nn[3] = {i % s.nx, (i / s.nx) % s.ny, i / (s.nx * s.ny)};
      double x[3];
      for (int d = 0; d < 3; d++) {
        x[d] = s.dx[d] * nn[d];
        if (x[d] > s.L[d] / 2.0) x[d] -= s.L[d];
        else if (x[d] <= -s.L[d] / 2.0) x[d] += s.L[d];
      }
      if (p.ic_flag[0] < -1) {
        if (rho_np[i].real() > 0.5) {
          wabm[i] = 0.0;
        } else {
          cd v = 1.0;
          const cd ph = s.chi_ab != 0 ? (s.neg_ab ? I : cd(1.0)) : cd(1.0);
          for (int j = 0; j < 2; j++) {
            const int dir = std::max(0, std::min(2, int(p.ic_dir[j])));
            const double arg = 2.0 * kPi * p.ic_period[j] * x[dir] / s.L[dir];
            if (p.ic_flag[j] == -2) v *= p.ic_pre[j] * std::tanh(std::cos(arg) / 0.2) * ph;
            else if (p.ic_flag[j] == -3) v *= p.ic_pre[j] * std::tanh(std::sin(arg) / 0.2) * ph;
          }
          wabm[i] = v;
        }
      } else if (p.ic_flag[0] == -1) {
        if (s.chi_ab != 0) {
          wabp[i] = p.ic_pre[0] * (rng.next() - 0.5) * (s.neg_ab ? cd(1.0) : I);
          wabm[i] = p.ic_pre[0] * (rng.next() - 0.5) * (s.neg_ab ? I : cd(1.0));
        }
        if (s.chi_bc != 0) {
          wbcp[i] = p.ic_pre[0] * (rng.next() - 0.5) * (s.neg_bc ? cd(1.0) : I);
          wbcm[i] = p.ic_pre[0] * (rng.next() - 0.5) * (s.neg_bc ? I : cd(1.0));
        }
        if (s.chi_ac != 0) {
          wacp[i] = p.ic_pre[0] * (rng.next() - 0.5) * (s.neg_ac ? cd(1.0) : I);
          wacm[i] = p.ic_pre[0] * (rng.next() - 0.5) * (s.neg_ac ? I : cd(1.0));
        }
        wpl[i] = p.ic_pre[0] * (rng.next() - 0.5) * I;
      }
    }
  }

  // device allocations
  const size_t m16 = (size_t)M * 16, m8 = (size_t)M * 8;
  const size_t pe = c->esz;
  SysDev& d = s.dev;
  memset(&d, 0, sizeof d);
  void* p_;
  int rc;
#define ALLOC(field, bytes)                                 \
  do {                                                      \
    rc = dalloc(c, s.allocs, &p_, (bytes));                 \
    if (rc) return rc;                                      \
    d.field = (decltype(d.field))p_;                        \
  } while (0)
  ALLOC(wpl, m16);
  ALLOC(wabp, m16);
  ALLOC(wabm, m16);
  ALLOC(rhoA, m16);
  ALLOC(rhoB, m16);
  ALLOC(wa, m16);
  ALLOC(wb, m16);
  ALLOC(rho_np, m16);
  ALLOC(exp_nrH, m16);
  ALLOC(hhat_k, m8);
  ALLOC(Xk, m8);
  ALLOC(X2k, m8);
  if (s.pairs) {
    ALLOC(wbcp, m16);
    ALLOC(wbcm, m16);
    ALLOC(wacp, m16);
    ALLOC(wacm, m16);
    ALLOC(XAk, m8);
    ALLOC(XBk, m8);
    ALLOC(XGk, m8);
  }
  // graft sites exist for every system with rod plates (code/initialize.cpp:
  // 160-205 computes them whether or not chains are grafted)
  if (s.has_sites) ALLOC(expl, m16);
  if (s.grafted) {
    ALLOC(wg, m16);
    ALLOC(rhog, m16);
    ALLOC(qg, (size_t)p.Ng * M * pe);
    ALLOC(qgd, (size_t)p.Ng * M * pe);
    ALLOC(expG, (size_t)M * pe);
  }
  if (c->ckpt) {
    // windows of K + 1 slices and the checkpoints; q_0 and q-dagger_0 are
    // written to slice 1 of their windows
    if (p.Nda == 0 || p.Nda == N)
      return fail(SCFT_EUNSUPPORTED, "checkpointed propagators need 0 < Nda < N");
    if ((M * pe) % 16 != 0) return fail(SCFT_EUNSUPPORTED, "checkpointed propagators need an even grid size");
    const int K = (int)std::ceil(std::sqrt((double)N));
    const int ne = (N + K - 1) / K;
    ALLOC(ck_sq, (size_t)(K + 1) * M * pe);
    ALLOC(ck_tq, (size_t)(K + 1) * M * pe);
    ALLOC(ck_cq, (size_t)ne * M * pe);
    d.q = (char*)d.ck_sq + (size_t)M * pe;
    d.qd = (char*)d.ck_tq + (size_t)M * pe;
    d.qlast = (char*)d.ck_sq + (size_t)std::min(K, N) * M * pe;
  } else {
    ALLOC(q, (size_t)N * M * pe);
    if (p.Nda != 0 && p.Nda != N) {
      ALLOC(qd, (size_t)N * M * pe);
    } else {
      d.qd = d.q;
    }
    d.qlast = (char*)d.q + (size_t)(N - 1) * M * pe;
  }
  ALLOC(expA, (size_t)M * pe);
  ALLOC(expB, (size_t)M * pe);
  if (p.do_CL) {
    ALLOC(etap, 2 * m16);
    ALLOC(etam, m16);
    ALLOC(rng_ctr, sizeof(unsigned long long));
    ALLOC(avgA, m16);
    ALLOC(avgB, m16);
    ALLOC(n_samples, sizeof(double));
  }
  ALLOC(red, scft::RED_NSLOTS * sizeof(double2));
  d.nblk = std::min(1024, (M + 255) / 256);
  ALLOC(partials, (size_t)scft::RED_NSLOTS * d.nblk * sizeof(double2));
  ALLOC(grad_partials, (size_t)8 * d.nblk * sizeof(double2));
  rc = dalloc(c, s.allocs, &p_, 8 * sizeof(double2));
  if (rc) return rc;
  s.grad_out = (double2*)p_;
  d.n_plates = std::min(p.n_exp_nr, 2);
  d.np_type = p.np_type;
  for (int k = 0; k < 3; k++) {
    d.Lb[k] = s.L[k];
    d.dxb[k] = s.dx[k];
  }
  d.len_nr = p.L_nr;
  d.R_nr = p.R_nr;
  d.xi_nr = p.xi_nr;
  for (int np = 0; np < 2; np++) set_plate_row(s, np);
  rc = dalloc(c, s.allocs, &p_, m16);
  if (rc) return rc;
  c->scratch[si] = p_;
#undef ALLOC
  HIPCK(hipMemset(d.red, 0, scft::RED_NSLOTS * sizeof(double2)));
  d.M = M;
  d.N = N;
  d.Na = p.Nda;
  d.nx = s.nx;
  d.ny = s.ny;
  d.nz = s.nz;
  d.dim = p.dim;
  d.neg_ab = s.neg_ab;
  d.do_cl = p.do_CL;
  d.np_chem = p.np_chem;
  d.nD = nD;
  d.V = V;
  d.dV = s.dV;
  d.rho0 = p.rho0;
  d.kappa = p.kappa;
  d.chi_ab = s.chi_ab;
  d.lam_imag = p.lam_imag;
  d.lam_real = p.lam_real;
  d.fA = fD;
  d.seed = p.seed;
  {
    // generate_1s_noise scales, code/1s_update.cpp:360-362; the reference
    // draws the w+ and AB- noise with lam_imag and the AB+ noise with
    // lam_real (code/1s_update.cpp:55, 91-92)
    const double lams[3] = {p.lam_imag, p.lam_real, p.lam_imag};
    for (int f = 0; f < 3; f++) {
      double sc = std::sqrt(2.0 * lams[f]);
      for (int dd = 0; dd < p.dim; dd++) sc /= std::sqrt(s.dx[dd]);
      d.noise_scale[f] = sc;
    }
  }
  d.sample_wait = p.sample_wait;
  d.sample_freq = p.sample_freq;
  if (p.do_CL) {
    HIPCK(hipMemset(d.rng_ctr, 0, sizeof(unsigned long long)));
    HIPCK(hipMemset(d.avgA, 0, m16));
    HIPCK(hipMemset(d.avgB, 0, m16));
    HIPCK(hipMemset(d.n_samples, 0, sizeof(double)));
  }
  HIPCK(hipMemcpy(d.rho_np, rho_np.data(), m16, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(d.hhat_k, hk.data(), m8, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(d.Xk, X.data(), m8, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(d.X2k, X2.data(), m8, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(d.wpl, wpl.data(), m16, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(d.wabp, wabp.data(), m16, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(d.wabm, wabm.data(), m16, hipMemcpyHostToDevice));
  HIPCK(hipMemset(d.rhoA, 0, m16));
  HIPCK(hipMemset(d.rhoB, 0, m16));
  d.neg_bc = s.neg_bc;
  d.neg_ac = s.neg_ac;
  d.chi_bc = s.chi_bc;
  d.chi_ac = s.chi_ac;
  d.grafted = s.grafted;
  d.Ng = s.grafted ? p.Ng : 0;
  d.n_exp = p.n_exp_nr;
  d.ng_per_np = ng_per_np;
  if (s.pairs) {
    HIPCK(hipMemcpy(d.wbcp, wbcp.data(), m16, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(d.wbcm, wbcm.data(), m16, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(d.wacp, wacp.data(), m16, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(d.wacm, wacm.data(), m16, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(d.XAk, XA.data(), m8, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(d.XBk, XB.data(), m8, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(d.XGk, XG.data(), m8, hipMemcpyHostToDevice));
  }
  if (s.has_sites) HIPCK(hipMemset(d.expl, 0, m16));
  if (s.grafted) HIPCK(hipMemset(d.rhog, 0, m16));

  // 1D tables along each axis: bond exp(-k^2/6)/n, smear exp(-k^2 a^2/2)/n
  for (int ax = 0; ax < 3; ax++) {
    const int n = ax == 0 ? s.nx : (ax == 1 ? s.ny : s.nz);
    std::vector<double> b(n), h(n);
    std::vector<float> bf(n);
    for (int m = 0; m < n; m++) {
      const double k = ax < p.dim ? kval(m, n, s.L[ax]) : 0.0;
      b[m] = std::exp(-k * k / 6.0) / n;
      h[m] = std::exp(-k * k * a2 / 2.0) / n;
      bf[m] = (float)b[m];
    }
    rc = dalloc(c, s.allocs, &p_, n * pe / 2);
    if (rc) return rc;
    s.bond_g[ax] = p_;
    if (c->prec == 0) HIPCK(hipMemcpy(p_, b.data(), n * 8, hipMemcpyHostToDevice));
    else HIPCK(hipMemcpy(p_, bf.data(), n * 4, hipMemcpyHostToDevice));
    rc = dalloc(c, s.allocs, &p_, n * 8);
    if (rc) return rc;
    s.smear_g[ax] = p_;
    HIPCK(hipMemcpy(p_, h.data(), n * 8, hipMemcpyHostToDevice));
  }
  return SCFT_OK;
}

int build_twiddles(scft_ctx* c) {
  std::vector<int> ns;
  for (auto& s : c->sys)
    for (int n : {s.nx, s.ny, s.nz})
      if (std::find(ns.begin(), ns.end(), n) == ns.end()) ns.push_back(n);
  size_t tot = 0;
  for (int n : ns) {
    c->tw_off[n] = tot;
    tot += n;
  }
  std::vector<double2> t64(tot);
  std::vector<float2> t32(tot);
  for (int n : ns) {
    const size_t o = c->tw_off[n];
    for (int m = 0; m < n; m++) {
      // exp(-2 pi i m / n) from long double sin/cos
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / n;
      t64[o + m] = make_double2((double)cosl(a), (double)sinl(a));
      t32[o + m] = make_float2((float)cosl(a), (float)sinl(a));
    }
  }
  int rc = dalloc(c, c->allocs, &c->tw64, tot * sizeof(double2));
  if (rc) return rc;
  rc = dalloc(c, c->allocs, &c->tw32, tot * sizeof(float2));
  if (rc) return rc;
  HIPCK(hipMemcpy(c->tw64, t64.data(), tot * sizeof(double2), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(c->tw32, t32.data(), tot * sizeof(float2), hipMemcpyHostToDevice));
  return SCFT_OK;
}

JobSpec job(int sys, int prec, void* in, void* out, int gkind, double scale) {
  JobSpec j{};
  j.sys = sys;
  j.prec = prec;
  j.in = in;
  j.out = out;
  j.gkind = gkind;
  j.scale = scale;
  return j;
}

int build_programs(scft_ctx* c) {
  const int ns = (int)c->sys.size();
  std::vector<JobSpec> smear, prop, prop_q, prop_qd, fwd, inv, noise, graft_q, graft_qd;
  for (int i = 0; i < ns; i++) {
    SysHost& s = c->sys[i];
    SysDev& d = s.dev;
    smear.push_back(job(i, 0, d.wa, d.wa, 2, 1.0));
    smear.push_back(job(i, 0, d.wb, d.wb, 2, 1.0));
    if (s.grafted) {
      smear.push_back(job(i, 0, d.wg, d.wg, 2, 1.0));
      // graft chains, code/grafted_nanoparticle.cpp:29-38, 60-71:
      // q[j] = exp(-smwg) IFFT[bond FFT q[j-1]], j = 1..Ng-1
      JobSpec g = job(i, c->prec, d.qg, d.qg, 1, 1.0);
      g.in_step = g.out_step = s.M;
      g.in_soff = -1;
      g.out_soff = 0;
      g.post_a = g.post_b = d.expG;
      g.post_switch = 0;
      g.steps_active = s.p.Ng;
      graft_q.push_back(g);
      g.in = g.out = d.qgd;
      graft_qd.push_back(g);
    }
    // propagator step: q[step-1] -> q[step], exp(-w) on the last store
    JobSpec q = job(i, c->prec, d.q, d.q, 1, 1.0);
    q.in_step = q.out_step = s.M;
    q.in_soff = -1;
    q.out_soff = 0;
    q.post_a = d.expA;
    q.post_b = d.expB;
    q.post_switch = s.Na > 0 ? s.Na : 0;
    q.steps_active = s.N;
    prop.push_back(q);
    prop_q.push_back(q);
    if (d.qd != d.q) {
      JobSpec qd = q;
      qd.in = qd.out = d.qd;
      qd.post_a = d.expB;
      qd.post_b = d.expA;
      qd.post_switch = s.N - s.Na;
      prop.push_back(qd);
      prop_qd.push_back(qd);
    }
    const double inv_m = 1.0 / double(s.M);
    for (void* a : {(void*)d.rhoA, (void*)d.rhoB, (void*)d.wpl, (void*)d.wabm, (void*)d.wabp})
      fwd.push_back(job(i, 0, a, a, 0, inv_m));
    for (void* a : {(void*)d.wpl, (void*)d.wabm, (void*)d.wabp}) inv.push_back(job(i, 0, a, a, 0, 1.0));
    // graft density and the pairs with chi != 0 (code/1s_update.cpp:27-52, 331-349)
    if (s.grafted) fwd.push_back(job(i, 0, d.rhog, d.rhog, 0, inv_m));
    if (s.pairs && s.chi_bc != 0.0)
      for (void* a : {(void*)d.wbcm, (void*)d.wbcp}) {
        fwd.push_back(job(i, 0, a, a, 0, inv_m));
        inv.push_back(job(i, 0, a, a, 0, 1.0));
      }
    if (s.pairs && s.chi_ac != 0.0)
      for (void* a : {(void*)d.wacm, (void*)d.wacp}) {
        fwd.push_back(job(i, 0, a, a, 0, inv_m));
        inv.push_back(job(i, 0, a, a, 0, 1.0));
      }
    if (d.etap) {
      noise.push_back(job(i, 0, d.etap, d.etap, 0, inv_m));
      noise.push_back(job(i, 0, (char*)d.etap + (size_t)s.M * 16, (char*)d.etap + (size_t)s.M * 16, 0, inv_m));
      noise.push_back(job(i, 0, d.etam, d.etam, 0, inv_m));
    }
  }
  int rc;
  if ((rc = build_program(c, scft::MODE_CONV, smear, &c->prog_smear))) return rc;
  if (!c->ckpt && (rc = build_program(c, scft::MODE_CONV, prop, &c->prog_prop))) return rc;
  c->any_graft = !graft_q.empty();
  for (auto& s : c->sys) c->any_pairs = c->any_pairs || s.pairs;
  if (c->any_graft) {
    for (auto& s : c->sys) c->maxNg = std::max(c->maxNg, s.grafted ? s.p.Ng : 0);
    if ((rc = build_program(c, scft::MODE_CONV, graft_q, &c->prog_graft_q))) return rc;
    if ((rc = build_program(c, scft::MODE_CONV, graft_qd, &c->prog_graft_qd))) return rc;
  }
  if (c->ckpt) {
    // chain steps inside the windows: slice u from slice u - 1
    std::vector<JobSpec> ckq, ckqd;
    for (int i = 0; i < ns; i++) {
      SysHost& s = c->sys[i];
      SysDev& d = s.dev;
      JobSpec q = job(i, c->prec, d.ck_sq, d.ck_sq, 1, 1.0);
      q.in_step = q.out_step = s.M;
      q.in_soff = -1;
      q.out_soff = 0;
      q.post_a = d.expA;
      q.post_b = d.expB;
      q.post_switch = s.Na;
      q.steps_active = 0;
      ckq.push_back(q);
      JobSpec qd = q;
      qd.in = qd.out = d.ck_tq;
      qd.post_a = d.expB;
      qd.post_b = d.expA;
      qd.post_switch = s.N - s.Na;
      ckqd.push_back(qd);
    }
    if ((rc = build_program(c, scft::MODE_CONV, ckq, &c->prog_ck_q))) return rc;
    if ((rc = build_program(c, scft::MODE_CONV, ckqd, &c->prog_ck_qd))) return rc;
  }
  c->split_chains = !prop_qd.empty() && !getenv("SCFT_ONE_STREAM") && !c->ckpt;
  if (c->split_chains) {
    if ((rc = build_program(c, scft::MODE_CONV, prop_q, &c->prog_prop_q))) return rc;
    if ((rc = build_program(c, scft::MODE_CONV, prop_qd, &c->prog_prop_qd))) return rc;
  }
  if ((rc = build_program(c, scft::MODE_FWD, fwd, &c->prog_fwd))) return rc;
  if ((rc = build_program(c, scft::MODE_INV, inv, &c->prog_inv))) return rc;
  if (!noise.empty() && (rc = build_program(c, scft::MODE_FWD, noise, &c->prog_noise))) return rc;
  c->any_cl = !noise.empty();
  c->prog_fft_fwd.resize(ns);
  c->prog_fft_inv.resize(ns);
  for (int i = 0; i < ns; i++) {
    std::vector<JobSpec> f = {job(i, 0, c->scratch[i], c->scratch[i], 0, 1.0 / double(c->sys[i].M))};
    std::vector<JobSpec> b = {job(i, 0, c->scratch[i], c->scratch[i], 0, 1.0)};
    if ((rc = build_program(c, scft::MODE_FWD, f, &c->prog_fft_fwd[i]))) return rc;
    if ((rc = build_program(c, scft::MODE_INV, b, &c->prog_fft_inv[i]))) return rc;
  }
  return SCFT_OK;
}

// ---------------------------------------------------------------------------
// stages
// ---------------------------------------------------------------------------
// Chains grafted to the plates, code/calc_poly_density.cpp:131-133, 155-160:
// free-end propagator, complementary propagator from the graft sites,
// graft density and the log-average Qga_exp
int stage_graft(scft_ctx* c, hipStream_t st) {
  if (!c->any_graft) return SCFT_OK;
  const int ns = (int)c->sys.size();
  int rc;
  for (int step = 1; step < c->maxNg; step++)
    if ((rc = run_program(c, c->prog_graft_q, step, st))) return rc;
  HIPCK(scft::launch_graft_init(c->prec, c->d_sys, ns, c->maxM, st));
  for (int step = 1; step < c->maxNg; step++)
    if ((rc = run_program(c, c->prog_graft_qd, step, st))) return rc;
  HIPCK(scft::launch_graft_density(c->prec, c->d_sys, ns, c->maxM, st));
  HIPCK(scft::launch_reduce(scft::RED_QGA, c->prec, c->d_sys, ns, c->nblk, st));
  return SCFT_OK;
}

// Checkpointed propagators and densities.  Array ids of launch_ck_copy: 0 q
// window, 1 q-dagger window, 2 checkpoints.  Slice u of a window holds the
// chain at contour index first + u - 1 of its segment.
int stage_ck_density(scft_ctx* c) {
  const int ns = (int)c->sys.size();
  const int K = c->ck_K, ne = c->ck_ne, esz = (int)c->esz;
  int rc = SCFT_OK;
  auto copy = [&](hipStream_t st, int src, int ss, int dst, int ds, int dst2, int ds2) -> int {
    HIPCK(scft::launch_ck_copy(c->d_sys, ns, c->maxM, esz, src, ss, dst, ds, dst2, ds2, st));
    return SCFT_OK;
  };
  auto steps = [&](const Program& p, hipStream_t st, int u0, int n, int first) -> int {
    for (int u = u0; u <= n; u++)
      if ((rc = run_program(c, p, u, st, first - 1))) return rc;
    return SCFT_OK;
  };
  // forward q pass from the bottom segment (q_0 in slice 1, saved as the
  // last checkpoint); each segment's top value is carried into slice 0 and
  // kept as the checkpoint of the segment above
  for (int e = ne - 1; e >= 0; e--) {
    const int lo = c->ck_lo[e], n = c->ck_n[e];
    if (e == ne - 1) {
      if ((rc = copy(c->st, 0, 1, 2, e, -1, 0)) || (rc = steps(c->prog_ck_q, c->st, 2, n, lo))) return rc;
    } else if ((rc = steps(c->prog_ck_q, c->st, 1, n, lo))) {
      return rc;
    }
    if (e > 0 && (rc = copy(c->st, 0, n, 0, 0, 2, e - 1))) return rc;
  }
  HIPCK(scft::launch_reduce(scft::RED_Q, c->prec, c->d_sys, ns, c->nblk, c->st));
  // density pass from the top segment: the q segment is recomputed from its
  // checkpoint (stream 1) while q-dagger advances by the same number of
  // steps (stream 2); the segment's density sum follows both
  for (int e = 0; e < ne; e++) {
    const int lo = c->ck_lo[e], n = c->ck_n[e];
    HIPCK(hipEventRecord(c->ck_fork[e], c->st));
    HIPCK(hipStreamWaitEvent(c->st2, c->ck_fork[e], 0));
    if (e == ne - 1 && e > 0) {
      // the bottom segment restarts from q_0 in slice 1
      if ((rc = copy(c->st, 2, e, 0, 1, -1, 0)) || (rc = steps(c->prog_ck_q, c->st, 2, n, lo))) return rc;
    } else if (e > 0) {
      if ((rc = copy(c->st, 2, e, 0, 0, -1, 0)) || (rc = steps(c->prog_ck_q, c->st, 1, n, lo))) return rc;
    }
    if (e == 0) {
      if ((rc = steps(c->prog_ck_qd, c->st2, 2, n, 0))) return rc;
    } else if ((rc = copy(c->st2, 1, K, 1, 0, -1, 0)) || (rc = steps(c->prog_ck_qd, c->st2, 1, n, e * K))) {
      return rc;
    }
    HIPCK(hipEventRecord(c->ck_join[e], c->st2));
    HIPCK(hipStreamWaitEvent(c->st, c->ck_join[e], 0));
    HIPCK(scft::launch_density_seg(c->prec, c->d_sys, ns, c->maxM, lo, n, e * K, e == 0, e == ne - 1, c->st));
  }
  return SCFT_OK;
}

int stage_poly_density(scft_ctx* c) {
  const int ns = (int)c->sys.size();
  HIPCK(scft::launch_assemble(c->d_sys, ns, c->maxM, c->st));
  int rc = run_program(c, c->prog_smear, 0);
  if (rc) return rc;
  if (c->prec == 1) HIPCK(scft::launch_reduce(scft::RED_SHIFT, 0, c->d_sys, ns, c->nblk, c->st));
  HIPCK(scft::launch_exp_prop(c->prec, c->d_sys, ns, c->maxM, c->st));
  if (c->ckpt) {
    if ((rc = stage_ck_density(c))) return rc;
    return stage_graft(c, c->st);
  }
  if (c->split_chains) {
    // graph branches: q on st, q-dagger on st2, the graft chains on st3
    // (they need only the smeared graft field; their reduction has its own
    // slot of the partials)
    HIPCK(hipEventRecord(c->ev_fork, c->st));
    HIPCK(hipStreamWaitEvent(c->st2, c->ev_fork, 0));
    if (c->any_graft) {
      HIPCK(hipStreamWaitEvent(c->st3, c->ev_fork, 0));
      if ((rc = stage_graft(c, c->st3))) return rc;
      HIPCK(hipEventRecord(c->ev_gjoin, c->st3));
    }
    for (int step = 1; step < c->maxN; step++) {
      if ((rc = run_program(c, c->prog_prop_q, step, c->st))) return rc;
      if ((rc = run_program(c, c->prog_prop_qd, step, c->st2))) return rc;
    }
    HIPCK(hipEventRecord(c->ev_join, c->st2));
    HIPCK(hipStreamWaitEvent(c->st, c->ev_join, 0));
    if (c->any_graft) HIPCK(hipStreamWaitEvent(c->st, c->ev_gjoin, 0));
    HIPCK(scft::launch_reduce(scft::RED_Q, c->prec, c->d_sys, ns, c->nblk, c->st));
    HIPCK(scft::launch_density(c->prec, c->d_sys, ns, c->maxM, c->st));
    return SCFT_OK;
  } else {
    for (int step = 1; step < c->maxN; step++) {
      rc = run_program(c, c->prog_prop, step);
      if (rc) return rc;
    }
  }
  HIPCK(scft::launch_reduce(scft::RED_Q, c->prec, c->d_sys, ns, c->nblk, c->st));
  HIPCK(scft::launch_density(c->prec, c->d_sys, ns, c->maxM, c->st));
  return stage_graft(c, c->st);
}

int stage_update(scft_ctx* c, int scheme) {
  int rc;
  if (c->any_cl) {
    // noise for this iteration (real space), then to k-space like the fields
    HIPCK(scft::launch_noise(c->d_sys, (int)c->sys.size(), c->maxM, c->st));
    if ((rc = run_program(c, c->prog_noise, 0))) return rc;
    HIPCK(scft::launch_noise_advance(c->d_sys, (int)c->sys.size(), c->st));
  }
  rc = run_program(c, c->prog_fwd, 0);
  if (rc) return rc;
  HIPCK(scft::launch_update(scheme, c->d_sys, (int)c->sys.size(), c->maxM, c->st));
  rc = run_program(c, c->prog_inv, 0);
  if (rc) return rc;
  return stage_poly_density(c);
}

// Graft sites of system si, code/initialize.cpp:160-196: the sum over plates
// of |grad rho| of the plate shape inflated by 1.05 in length and radius
// (spectral derivative, field_gradient at code/array_utils.cpp:565-579),
// normalised to unit integral.
int graft_sites(scft_ctx* c, int si) {
  SysHost& s = c->sys[si];
  const int M = s.M;
  const size_t m16 = (size_t)M * 16;
  std::vector<void*> tmp;
  auto scratch_alloc = [&](void** p) -> int {
    if (hipMalloc(p, m16) != hipSuccess) return fail(SCFT_ENOMEM, "no memory for the graft sites");
    tmp.push_back(*p);
    return SCFT_OK;
  };
  void *spec = nullptr, *g2 = nullptr;
  int rc = scratch_alloc(&spec);
  if (rc == SCFT_OK) rc = scratch_alloc(&g2);
  // everything runs on the context stream, behind the work that still reads
  // the FFT scratch; the host shape buffer is refilled only after the stream
  // has drained its previous upload
  std::vector<cd> shape(M);
  for (int np = 0; np < std::min(s.p.n_exp_nr, 2) && rc == SCFT_OK; np++) {
    if (hipStreamSynchronize(c->st) != hipSuccess) {
      rc = fail(SCFT_EDEVICE, "stream failed");
      break;
    }
    for (int i = 0; i < M; i++) {
      const int nn[3] = {i % s.nx, (i / s.nx) % s.ny, i / (s.nx * s.ny)};
      shape[i] = plate_rho(s, np, nn, 1.05);
    }
    if (hipMemcpyAsync(c->scratch[si], shape.data(), m16, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipMemsetAsync(g2, 0, m16, c->st) != hipSuccess) {
      rc = fail(SCFT_EDEVICE, "copy failed");
      break;
    }
    if ((rc = run_program(c, c->prog_fft_fwd[si], 0))) break;
    if (hipMemcpyAsync(spec, c->scratch[si], m16, hipMemcpyDeviceToDevice, c->st) != hipSuccess) {
      rc = fail(SCFT_EDEVICE, "copy failed");
      break;
    }
    for (int d = 0; d < s.p.dim && rc == SCFT_OK; d++) {
      if (scft::launch_ik_mult(c->d_sys + si, M, (const double2*)spec, (double2*)c->scratch[si], d, c->st) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "graft-site kernel failed");
      if (rc == SCFT_OK) rc = run_program(c, c->prog_fft_inv[si], 0);
      if (rc == SCFT_OK &&
          scft::launch_sites_accum(M, (const double2*)c->scratch[si], (double2*)g2, 0, c->st) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "graft-site kernel failed");
    }
    if (rc == SCFT_OK && scft::launch_sites_accum(M, (const double2*)g2, s.dev.expl, 1, c->st) != hipSuccess)
      rc = fail(SCFT_EDEVICE, "graft-site kernel failed");
  }
  // normalise on the host (set-up only)
  if (rc == SCFT_OK) {
    std::vector<cd> e(M);
    if (hipMemcpyAsync(e.data(), s.dev.expl, m16, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess) {
      rc = fail(SCFT_EDEVICE, "copy failed");
    } else {
      cd total = 0.0;
      for (int i = 0; i < M; i++) total += e[i];
      total *= s.dV;
      s.expl_Z = total;
      cd after = 0.0;
      for (int i = 0; i < M; i++) {
        e[i] /= total;
        after += e[i];
      }
      s.expl_total = after * s.dV;
      if (hipMemcpyAsync(s.dev.expl, e.data(), m16, hipMemcpyHostToDevice, c->st) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "copy failed");
    }
  }
  (void)hipStreamSynchronize(c->st);
  for (void* p : tmp) (void)hipFree(p);
  return rc;
}

// Plate derivatives of the graft term -n ng_per_np Re Qga, Qga =
// integ(expl_grafts L) with L = log q_g[Ng-1] (code/calc_h.cpp:31-35), at
// fixed fields.  expl_grafts = sum_p g_p / Z, g_p = |grad shape_p| of the
// inflated plate shape (spectral gradient, as graft_sites computes it), Z =
// integ(sum_p g_p), so for a parameter t of plate p
//   d Qga / dt = (dV / Z) sum_r (d g_p / dt) (L - Qga),
//   d g_p / dt = sum_d G_d G'_d / g_p,  G_d = D_d shape_p, G'_d = D_d d shape_p / dt,
// the exact derivative of the discrete graft sites.  out[4 p + k]: Re part of
// dH/dc_k (k < 3) and dH/dtheta (k = 3) about `axis`.
int graft_site_grad(scft_ctx* c, int si, const double* axis, cd Qga, double* out) {
  SysHost& s = c->sys[si];
  const int M = s.M, dim = s.p.dim;
  const size_t m16 = (size_t)M * 16;
  const SysDev* row = c->d_sys + si;
  std::vector<void*> tmp;
  auto get = [&](double2** p) -> int {
    if (hipMalloc((void**)p, m16) != hipSuccess) return fail(SCFT_ENOMEM, "no memory for the graft-site derivative");
    tmp.push_back(*p);
    return SCFT_OK;
  };
  double2 *spec = nullptr, *G[3] = {nullptr, nullptr, nullptr}, *g2 = nullptr, *gq = nullptr, *S = nullptr,
          *part = nullptr;
  int rc = get(&spec);
  for (int d = 0; d < dim && rc == SCFT_OK; d++) rc = get(&G[d]);
  if (rc == SCFT_OK) rc = get(&g2);
  if (rc == SCFT_OK) rc = get(&gq);
  if (rc == SCFT_OK) rc = get(&S);
  const int nblk = s.dev.nblk;
  if (rc == SCFT_OK && hipMalloc((void**)&part, (size_t)nblk * sizeof(double2)) != hipSuccess)
    rc = fail(SCFT_ENOMEM, "no memory for the graft-site derivative");
  if (part) tmp.push_back(part);
  double2* scr = (double2*)c->scratch[si];
  // D_d of the field in scr, into dst (the spectrum is kept in spec)
  auto grad_of_scr = [&](bool fwd, int d, double2* dst) -> int {
    int r;
    if (fwd) {
      if ((r = run_program(c, c->prog_fft_fwd[si], 0))) return r;
      HIPCK(hipMemcpyAsync(spec, scr, m16, hipMemcpyDeviceToDevice, c->st));
    }
    HIPCK(scft::launch_ik_mult(row, M, spec, scr, d, c->st));
    if ((r = run_program(c, c->prog_fft_inv[si], 0))) return r;
    HIPCK(hipMemcpyAsync(dst, scr, m16, hipMemcpyDeviceToDevice, c->st));
    return SCFT_OK;
  };
  const double pref = -double(s.p.n_exp_nr) * s.info.ng_per_np;
  for (int k = 0; k < 8; k++) out[k] = 0.0;
  std::vector<double2> h(nblk);
  for (int p = 0; p < s.dev.n_plates && rc == SCFT_OK; p++) {
    // G_d and g_p of the inflated shape
    if (scft::launch_site_shape(row, M, p, -1, axis, 1.05, scr, c->st) != hipSuccess ||
        hipMemsetAsync(g2, 0, m16, c->st) != hipSuccess || hipMemsetAsync(gq, 0, m16, c->st) != hipSuccess) {
      rc = fail(SCFT_EDEVICE, "graft-site derivative kernel failed");
      break;
    }
    for (int d = 0; d < dim && rc == SCFT_OK; d++) {
      if ((rc = grad_of_scr(d == 0, d, G[d]))) break;
      if (scft::launch_sites_accum(M, G[d], g2, 0, c->st) != hipSuccess) rc = fail(SCFT_EDEVICE, "kernel failed");
    }
    if (rc == SCFT_OK && scft::launch_sites_accum(M, g2, gq, 1, c->st) != hipSuccess)
      rc = fail(SCFT_EDEVICE, "kernel failed");
    // round-off zeros of g: |g| <= 1e-12 max |g|
    double g2_min = 0.0;
    if (rc == SCFT_OK) {
      std::vector<double2> gh(M);
      if (hipMemcpyAsync(gh.data(), gq, m16, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
          hipStreamSynchronize(c->st) != hipSuccess) {
        rc = fail(SCFT_EDEVICE, "copy failed");
        break;
      }
      double gmax2 = 0.0;
      for (int i = 0; i < M; i++) gmax2 = std::max(gmax2, gh[i].x * gh[i].x + gh[i].y * gh[i].y);
      g2_min = 1e-24 * gmax2;
    }
    for (int lam = 0; lam < 4 && rc == SCFT_OK; lam++) {
      if (lam >= dim && lam < 3) continue;  // no centre coordinate along an absent axis
      if (scft::launch_site_shape(row, M, p, lam, axis, 1.05, scr, c->st) != hipSuccess ||
          hipMemsetAsync(S, 0, m16, c->st) != hipSuccess) {
        rc = fail(SCFT_EDEVICE, "graft-site derivative kernel failed");
        break;
      }
      for (int d = 0; d < dim && rc == SCFT_OK; d++) {
        if ((rc = grad_of_scr(d == 0, d, g2))) break;  // g2 reused as G'_d
        if (scft::launch_cmul_acc(M, G[d], g2, S, c->st) != hipSuccess) rc = fail(SCFT_EDEVICE, "kernel failed");
      }
      if (rc) break;
      if (scft::launch_site_grad_partial(c->prec, row, nblk, S, gq, Qga.real(), Qga.imag(), g2_min, part, c->st) !=
          hipSuccess) {
        rc = fail(SCFT_EDEVICE, "kernel failed");
        break;
      }
      if (hipMemcpyAsync(h.data(), part, (size_t)nblk * sizeof(double2), hipMemcpyDeviceToHost, c->st) !=
              hipSuccess ||
          hipStreamSynchronize(c->st) != hipSuccess) {
        rc = fail(SCFT_EDEVICE, "copy failed");
        break;
      }
      cd sum = 0.0;
      for (int b = 0; b < nblk; b++) sum += cd(h[b].x, h[b].y);
      out[4 * p + lam] = pref * (sum * s.dV / s.expl_Z).real();
    }
  }
  (void)hipStreamSynchronize(c->st);
  for (void* q : tmp) (void)hipFree(q);
  return rc;
}

// one iteration of the reference main loop (code/simulate.cpp:86-92):
// update, then the complex-Langevin averages
int stage_iteration(scft_ctx* c, int scheme) {
  int rc = stage_update(c, scheme);
  if (rc) return rc;
  if (c->any_cl) HIPCK(scft::launch_cl_accumulate(c->d_sys, (int)c->sys.size(), c->maxM, c->st));
  return SCFT_OK;
}

// Qga_exp = integ(expl_grafts log q_g[Ng-1]) from its reduction; with the
// gauge shift c the free-end propagator carries exp(Ng c)
cd qga(const SysHost& s, const double2* red) {
  const cd sum = cd(red[scft::RED_QGA].x, red[scft::RED_QGA].y) * s.dV;
  return sum - double(s.p.Ng) * red[scft::RED_SHIFT].y * s.expl_total;
}

int check_ctx(scft_ctx* c, int sys) {
  if (!c) return fail(SCFT_EINVAL, "null context");
  if (sys < 0 || sys >= (int)c->sys.size()) return fail(SCFT_EINVAL, "system index out of range");
  HIPCK(hipSetDevice(c->device));
  return SCFT_OK;
}

int copy_out(scft_ctx* c, void* host, const void* dev, size_t bytes) {
  HIPCK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->st));
  HIPCK(hipStreamSynchronize(c->st));
  return SCFT_OK;
}

int copy_in(scft_ctx* c, void* dev, const void* host, size_t bytes) {
  HIPCK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->st));
  HIPCK(hipStreamSynchronize(c->st));
  return SCFT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* scft_last_error(void) { return g_err.c_str(); }

int scft_create(const scft_params* params, int n_systems, int device, int precision, scft_ctx** out) {
  if (!params || !out || n_systems < 1) return fail(SCFT_EINVAL, "bad arguments to scft_create");
  const bool ckpt = (precision & SCFT_OPT_CHECKPOINT) != 0;
  precision &= ~SCFT_OPT_CHECKPOINT;
  if (precision != SCFT_PREC_FP64 && precision != SCFT_PREC_FP32) return fail(SCFT_EINVAL, "bad precision");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SCFT_EDEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(SCFT_EINVAL, "device index out of range");
  HIPCK(hipSetDevice(device));
  scft_ctx* c = new scft_ctx();
  c->device = device;
  c->prec = precision;
  c->esz = precision == SCFT_PREC_FP64 ? 16 : 8;
  c->sys.resize(n_systems);
  c->scratch.resize(n_systems);
  c->ckpt = ckpt;
  int rc = SCFT_OK;
  if (ckpt) {
    // one segment plan for the context: every system has the same N
    const int N = params[0].Nda + params[0].Ndb;
    for (int i = 0; i < n_systems; i++)
      if (params[i].Nda + params[i].Ndb != N) {
        delete c;
        return fail(SCFT_EUNSUPPORTED, "checkpointed propagators need the same N in all systems");
      }
    c->ck_K = (int)std::ceil(std::sqrt((double)N));
    c->ck_ne = (N + c->ck_K - 1) / c->ck_K;
    for (int e = 0; e < c->ck_ne; e++) {
      const int hi = N - 1 - e * c->ck_K, lo = std::max(0, N - (e + 1) * c->ck_K);
      c->ck_lo.push_back(lo);
      c->ck_n.push_back(hi - lo + 1);
    }
    c->ck_fork.assign(c->ck_ne, nullptr);
    c->ck_join.assign(c->ck_ne, nullptr);
    for (int e = 0; e < c->ck_ne; e++)
      if (hipEventCreateWithFlags(&c->ck_fork[e], hipEventDisableTiming) != hipSuccess ||
          hipEventCreateWithFlags(&c->ck_join[e], hipEventDisableTiming) != hipSuccess) {
        scft_destroy(c);
        return fail(SCFT_EDEVICE, "hipEventCreate failed");
      }
  }
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->st3, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_gjoin, hipEventDisableTiming) != hipSuccess) {
    scft_destroy(c);
    return fail(SCFT_EDEVICE, "hipStreamCreate failed");
  }
  for (int i = 0; i < n_systems && rc == SCFT_OK; i++) {
    c->sys[i].p = params[i];
    rc = setup_system(c, i);
  }
  if (rc == SCFT_OK) rc = build_twiddles(c);
  if (rc == SCFT_OK) {
    for (auto& s : c->sys) {
      c->maxM = std::max(c->maxM, s.M);
      c->maxN = std::max(c->maxN, s.N);
      c->nblk = std::max(c->nblk, s.dev.nblk);
    }
    void* p;
    rc = dalloc(c, c->allocs, &p, n_systems * sizeof(SysDev));
    if (rc == SCFT_OK) {
      c->d_sys = (SysDev*)p;
      std::vector<SysDev> rows;
      for (auto& s : c->sys) rows.push_back(s.dev);
      if (hipMemcpy(c->d_sys, rows.data(), rows.size() * sizeof(SysDev), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "upload of system table failed");
    }
  }
  if (rc == SCFT_OK) rc = build_programs(c);
  // exp_nrH = forward FFT of the plate density (code/initialize.cpp:236-239)
  for (int i = 0; i < n_systems && rc == SCFT_OK; i++) {
    SysHost& s = c->sys[i];
    if (hipMemcpyAsync(c->scratch[i], s.dev.rho_np, (size_t)s.M * 16, hipMemcpyDeviceToDevice, c->st) != hipSuccess)
      rc = fail(SCFT_EDEVICE, "copy failed");
    if (rc == SCFT_OK) rc = run_program(c, c->prog_fft_fwd[i], 0);
    if (rc == SCFT_OK &&
        hipMemcpyAsync(s.dev.exp_nrH, c->scratch[i], (size_t)s.M * 16, hipMemcpyDeviceToDevice, c->st) != hipSuccess)
      rc = fail(SCFT_EDEVICE, "copy failed");
  }
  for (int i = 0; i < n_systems && rc == SCFT_OK; i++)
    if (c->sys[i].has_sites) rc = graft_sites(c, i);
  if (rc == SCFT_OK && hipStreamSynchronize(c->st) != hipSuccess) rc = fail(SCFT_EDEVICE, "set-up failed on the device");
  if (rc == SCFT_OK && !getenv("SCFT_NO_TUNE")) {
    // tuning runs on zeroed propagators and scratch fields (recomputed later)
    for (auto& s : c->sys) {
      if (rc == SCFT_OK && hipMemsetAsync(s.dev.q, 0, (size_t)2 * s.M * c->esz, c->st) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "memset failed");
      if (rc == SCFT_OK && s.dev.qd != s.dev.q &&
          hipMemsetAsync(s.dev.qd, 0, (size_t)2 * s.M * c->esz, c->st) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "memset failed");
      if (rc == SCFT_OK && hipMemsetAsync(s.dev.wa, 0, (size_t)s.M * 16, c->st) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "memset failed");
      if (rc == SCFT_OK && hipMemsetAsync(s.dev.wb, 0, (size_t)s.M * 16, c->st) != hipSuccess)
        rc = fail(SCFT_EDEVICE, "memset failed");
    }
    if (rc == SCFT_OK && !c->ckpt) rc = autotune(c, c->prog_prop, 1);
    if (rc == SCFT_OK && c->ckpt) rc = autotune(c, c->prog_ck_q, 2);
    if (rc == SCFT_OK && c->ckpt) rc = autotune(c, c->prog_ck_qd, 2);
    if (rc == SCFT_OK && c->split_chains) rc = autotune(c, c->prog_prop_q, 1);
    if (rc == SCFT_OK && c->split_chains) rc = autotune(c, c->prog_prop_qd, 1);
    if (rc == SCFT_OK) rc = autotune(c, c->prog_smear, 0);
    // the field programs work in place on w+, wAB+-, rhoA, rhoB: keep a copy
    std::vector<void*> saved;
    for (auto& s : c->sys) {
      void* p = nullptr;
      if (rc == SCFT_OK && hipMalloc(&p, (size_t)5 * s.M * 16) != hipSuccess)
        rc = fail(SCFT_ENOMEM, "no memory for the tuning copy");
      saved.push_back(p);
      void* src[5] = {s.dev.wpl, s.dev.wabp, s.dev.wabm, s.dev.rhoA, s.dev.rhoB};
      for (int a = 0; a < 5 && rc == SCFT_OK; a++)
        if (hipMemcpyAsync((char*)p + (size_t)a * s.M * 16, src[a], (size_t)s.M * 16, hipMemcpyDeviceToDevice,
                           c->st) != hipSuccess)
          rc = fail(SCFT_EDEVICE, "copy failed");
    }
    if (rc == SCFT_OK) rc = autotune(c, c->prog_fwd, 0);
    if (rc == SCFT_OK) rc = autotune(c, c->prog_inv, 0);
    for (size_t i = 0; i < c->sys.size(); i++) {
      auto& s = c->sys[i];
      void* dst[5] = {s.dev.wpl, s.dev.wabp, s.dev.wabm, s.dev.rhoA, s.dev.rhoB};
      for (int a = 0; a < 5 && rc == SCFT_OK && saved[i]; a++)
        if (hipMemcpyAsync(dst[a], (char*)saved[i] + (size_t)a * s.M * 16, (size_t)s.M * 16,
                           hipMemcpyDeviceToDevice, c->st) != hipSuccess)
          rc = fail(SCFT_EDEVICE, "copy failed");
    }
    (void)hipStreamSynchronize(c->st);
    for (void* p : saved)
      if (p) (void)hipFree(p);
    if (rc == SCFT_OK && hipStreamSynchronize(c->st) != hipSuccess) rc = fail(SCFT_EDEVICE, "autotune failed");
  }
  if (rc != SCFT_OK) {
    std::string msg = g_err;
    scft_destroy(c);
    g_err = msg;
    return rc;
  }
  *out = c;
  return SCFT_OK;
}

int scft_destroy(scft_ctx* c) {
  if (!c) return SCFT_OK;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->st2) (void)hipStreamSynchronize(c->st2);
  if (c->st3) (void)hipStreamSynchronize(c->st3);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  for (auto& s : c->sys)
    for (void* p : s.allocs) (void)hipFree(p);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->st) (void)hipStreamDestroy(c->st);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st3) (void)hipStreamDestroy(c->st3);
  for (hipEvent_t e : c->ck_fork)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->ck_join)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_gjoin) (void)hipEventDestroy(c->ev_gjoin);
  delete c;
  return SCFT_OK;
}

int scft_get_info(scft_ctx* c, int sys, scft_sysinfo* out) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!out) return fail(SCFT_EINVAL, "null output");
  *out = c->sys[sys].info;
  return SCFT_OK;
}

int scft_set_fields(scft_ctx* c, int sys, const double* wpl, const double* wabp, const double* wabm) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  SysHost& s = c->sys[sys];
  const size_t b = (size_t)s.M * 16;
  if (wpl && (rc = copy_in(c, s.dev.wpl, wpl, b))) return rc;
  if (wabp && (rc = copy_in(c, s.dev.wabp, wabp, b))) return rc;
  if (wabm && (rc = copy_in(c, s.dev.wabm, wabm, b))) return rc;
  return SCFT_OK;
}

int scft_get_fields(scft_ctx* c, int sys, double* wpl, double* wabp, double* wabm) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  SysHost& s = c->sys[sys];
  const size_t b = (size_t)s.M * 16;
  if (wpl && (rc = copy_out(c, wpl, s.dev.wpl, b))) return rc;
  if (wabp && (rc = copy_out(c, wabp, s.dev.wabp, b))) return rc;
  if (wabm && (rc = copy_out(c, wabm, s.dev.wabm, b))) return rc;
  return SCFT_OK;
}

int scft_get_density(scft_ctx* c, int sys, double* rhoA, double* rhoB) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  SysHost& s = c->sys[sys];
  const size_t b = (size_t)s.M * 16;
  if (rhoA && (rc = copy_out(c, rhoA, s.dev.rhoA, b))) return rc;
  if (rhoB && (rc = copy_out(c, rhoB, s.dev.rhoB, b))) return rc;
  return SCFT_OK;
}

int scft_get_smeared(scft_ctx* c, int sys, double* smwa, double* smwb) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  SysHost& s = c->sys[sys];
  const size_t b = (size_t)s.M * 16;
  if (smwa && (rc = copy_out(c, smwa, s.dev.wa, b))) return rc;
  if (smwb && (rc = copy_out(c, smwb, s.dev.wb, b))) return rc;
  return SCFT_OK;
}

int scft_get_cavity(scft_ctx* c, int sys, double* rho) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!rho) return fail(SCFT_EINVAL, "null output");
  SysHost& s = c->sys[sys];
  return copy_out(c, rho, s.dev.rho_np, (size_t)s.M * 16);
}

int scft_get_Q(scft_ctx* c, int sys, double* Q) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!Q) return fail(SCFT_EINVAL, "null output");
  SysHost& s = c->sys[sys];
  double2 red[scft::RED_NSLOTS];
  if ((rc = copy_out(c, red, s.dev.red, sizeof red))) return rc;
  const cd Qs = cd(red[scft::RED_Q].x, red[scft::RED_Q].y) * (s.dV / s.info.V);
  const cd Qt = Qs * std::exp(-double(s.N) * red[scft::RED_SHIFT].y);
  Q[0] = Qt.real();
  Q[1] = Qt.imag();
  return SCFT_OK;
}

int scft_get_noise(scft_ctx* c, int sys, double* a, double* b, double* m) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  SysHost& s = c->sys[sys];
  if (!s.dev.etap) return fail(SCFT_EINVAL, "system has no Langevin noise (do_CL = 0)");
  const size_t bytes = (size_t)s.M * 16;
  if (a && (rc = copy_out(c, a, s.dev.etap, bytes))) return rc;
  if (b && (rc = copy_out(c, b, (char*)s.dev.etap + bytes, bytes))) return rc;
  if (m && (rc = copy_out(c, m, s.dev.etam, bytes))) return rc;
  return SCFT_OK;
}

int scft_get_averages(scft_ctx* c, int sys, double* avg_a, double* avg_b, double* n_samples) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  SysHost& s = c->sys[sys];
  if (!s.dev.avgA) return fail(SCFT_EINVAL, "system has no Langevin averages (do_CL = 0)");
  double n = 0.0;
  if ((rc = copy_out(c, &n, s.dev.n_samples, sizeof n))) return rc;
  if (n_samples) *n_samples = n;
  const size_t bytes = (size_t)s.M * 16;
  double* outs[2] = {avg_a, avg_b};
  const void* src[2] = {s.dev.avgA, s.dev.avgB};
  for (int k = 0; k < 2; k++) {
    if (!outs[k]) continue;
    if ((rc = copy_out(c, outs[k], src[k], bytes))) return rc;
    // write_avg_data_bin divides the sums by n_samples (code/io_utils.cpp:368-369)
    if (n > 0.0)
      for (size_t i = 0; i < 2 * (size_t)s.M; i++) outs[k][i] /= n;
  }
  return SCFT_OK;
}

namespace {
// Named complex128 arrays of M values (the reference's globals,
// code/globals.h:76-123); nullptr when the system does not hold the array.
void* named_array(scft_ctx* c, SysHost& s, const char* name, bool writable) {
  const std::string n(name);
  SysDev& d = s.dev;
  if (n == "wpl") return d.wpl;
  if (n == "wabp") return d.wabp;
  if (n == "wabm") return d.wabm;
  if (n == "wbcp") return d.wbcp;
  if (n == "wbcm") return d.wbcm;
  if (n == "wacp") return d.wacp;
  if (n == "wacm") return d.wacm;
  if (writable) return nullptr;
  if (n == "rhoda") return d.rhoA;
  if (n == "rhodb") return d.rhoB;
  if (n == "smwa") return d.wa;
  if (n == "smwb") return d.wb;
  if (n == "rho_exp_nr") return d.rho_np;
  if (n == "smwg") return d.wg;
  if (n == "expl_grafts") return d.expl;
  if (n == "rhog_exp") return d.rhog;
  if (n.rfind("qg:", 0) == 0 && d.qg && c->prec == 0) {
    const int j = std::atoi(n.c_str() + 3);
    if (j >= 0 && j < s.p.Ng) return (char*)d.qg + (size_t)j * s.M * 16;
  }
  return nullptr;
}
}  // namespace

int scft_get_array(scft_ctx* c, int sys, const char* name, double* out) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!name || !out) return fail(SCFT_EINVAL, "null argument");
  SysHost& s = c->sys[sys];
  const void* p = named_array(c, s, name, false);
  if (!p) return fail(SCFT_EINVAL, std::string("system holds no array named ") + name);
  return copy_out(c, out, p, (size_t)s.M * 16);
}

int scft_set_array(scft_ctx* c, int sys, const char* name, const double* in) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!name || !in) return fail(SCFT_EINVAL, "null argument");
  SysHost& s = c->sys[sys];
  void* p = named_array(c, s, name, true);
  if (!p) return fail(SCFT_EINVAL, std::string("system holds no settable field named ") + name);
  return copy_in(c, p, in, (size_t)s.M * 16);
}

int scft_get_Qga(scft_ctx* c, int sys, double* Qga) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!Qga) return fail(SCFT_EINVAL, "null output");
  SysHost& s = c->sys[sys];
  if (!s.grafted) return fail(SCFT_EINVAL, "system has no grafted chains");
  double2 red[scft::RED_NSLOTS];
  if ((rc = copy_out(c, red, s.dev.red, sizeof red))) return rc;
  const cd q = qga(s, red);
  Qga[0] = q.real();
  Qga[1] = q.imag();
  return SCFT_OK;
}

int scft_set_plates(scft_ctx* c, int sys, const double* centres, const double* u) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!centres || !u) return fail(SCFT_EINVAL, "null argument");
  SysHost& s = c->sys[sys];
  const int np = s.dev.n_plates;
  if (np == 0) return fail(SCFT_EINVAL, "system has no plates");
  for (int j = 0; j < np; j++) {
    double m2 = 0.0;
    for (int d = 0; d < s.p.dim; d++) m2 += u[3 * j + d] * u[3 * j + d];
    if (m2 <= 0.0) return fail(SCFT_EINVAL, "plate orientation must be nonzero");
    for (int d = 0; d < s.p.dim; d++) {
      if (centres[3 * j + d] < 0.0 || centres[3 * j + d] > 1.0)
        return fail(SCFT_EINVAL, "plate centres are fractions of the box (0..1)");
      s.p.exp_nr_c[j][d] = centres[3 * j + d];
      s.p.exp_nr_u[j][d] = u[3 * j + d] / std::sqrt(m2);
    }
    set_plate_row(s, j);
  }
  SysDev* row = c->d_sys + sys;
  HIPCK(hipMemcpyAsync(row, &s.dev, sizeof(SysDev), hipMemcpyHostToDevice, c->st));
  HIPCK(scft::launch_cavity(row, 1, s.M, c->st));
  // exp_nrH = forward transform of the plate density
  HIPCK(hipMemcpyAsync(c->scratch[sys], s.dev.rho_np, (size_t)s.M * 16, hipMemcpyDeviceToDevice, c->st));
  if ((rc = run_program(c, c->prog_fft_fwd[sys], 0))) return rc;
  HIPCK(hipMemcpyAsync(s.dev.exp_nrH, c->scratch[sys], (size_t)s.M * 16, hipMemcpyDeviceToDevice, c->st));
  // volume bookkeeping, code/initialize.cpp:249-355
  HIPCK(scft::launch_reduce(scft::RED_VEXP, 0, row, 1, s.dev.nblk, c->st));
  double2 red[scft::RED_NSLOTS];
  if ((rc = copy_out(c, red, s.dev.red, sizeof red))) return rc;
  const double V_exp = red[scft::RED_VEXP].x * s.dV;
  scft_sysinfo& inf = s.info;
  const double phiD_old = inf.phiD;
  inf.np_frac = V_exp / inf.Vf;
  inf.V_1_exp_np = V_exp / double(s.p.n_exp_nr);
  inf.phiD = 1.0 - s.p.phiH - inf.np_frac - inf.phiG;
  inf.nD = inf.phiD * s.p.rho0 * inf.Vf / s.N;
  s.dev.nD = inf.nD;
  if (phiD_old != 0.0) {
    // X(k) and X2(k) carry phiD as a factor
    HIPCK(scft::launch_scale_real(s.dev.Xk, s.M, inf.phiD / phiD_old, c->st));
    HIPCK(scft::launch_scale_real(s.dev.X2k, s.M, inf.phiD / phiD_old, c->st));
    if (s.pairs) {
      HIPCK(scft::launch_scale_real(s.dev.XAk, s.M, inf.phiD / phiD_old, c->st));
      HIPCK(scft::launch_scale_real(s.dev.XBk, s.M, inf.phiD / phiD_old, c->st));
    }
  }
  HIPCK(hipMemcpyAsync(row, &s.dev, sizeof(SysDev), hipMemcpyHostToDevice, c->st));
  if (s.has_sites) {
    // the graft sites follow the plates
    HIPCK(hipMemsetAsync(s.dev.expl, 0, (size_t)s.M * 16, c->st));
    if ((rc = graft_sites(c, sys))) return rc;
  }
  HIPCK(hipStreamSynchronize(c->st));
  return SCFT_OK;
}

int scft_grad_plates(scft_ctx* c, int sys, const double* axis, double* out) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!axis || !out) return fail(SCFT_EINVAL, "null argument");
  SysHost& s = c->sys[sys];
  double2 red[scft::RED_NSLOTS];
  if ((rc = copy_out(c, red, s.dev.red, sizeof red))) return rc;
  const cd Qs = cd(red[scft::RED_Q].x, red[scft::RED_Q].y) * (s.dV / s.info.V);
  const cd logQ = std::log(Qs) - double(s.N) * red[scft::RED_SHIFT].y;
  HIPCK(scft::launch_plate_grad(c->d_sys + sys, s.dev.nblk, axis, logQ.real(), logQ.imag(), s.grad_out, c->st));
  double2 g[8];
  if ((rc = copy_out(c, g, s.grad_out, sizeof g))) return rc;
  for (int k = 0; k < 8; k++) {
    out[2 * k] = k < 4 * s.dev.n_plates ? g[k].x : 0.0;
    out[2 * k + 1] = k < 4 * s.dev.n_plates ? g[k].y : 0.0;
  }
  if (s.grafted) {
    double gg[8];
    if ((rc = graft_site_grad(c, sys, axis, qga(s, red), gg))) return rc;
    for (int k = 0; k < 4 * s.dev.n_plates; k++) out[2 * k] += gg[k];
  }
  return SCFT_OK;
}

int scft_fft(scft_ctx* c, int sys, double* inout, int dir) {
  int rc = check_ctx(c, sys);
  if (rc) return rc;
  if (!inout) return fail(SCFT_EINVAL, "null argument");
  if (dir != -1 && dir != 1) return fail(SCFT_EINVAL, "dir must be -1 or +1");
  SysHost& s = c->sys[sys];
  const size_t b = (size_t)s.M * 16;
  HIPCK(hipMemcpyAsync(c->scratch[sys], inout, b, hipMemcpyHostToDevice, c->st));
  rc = run_program(c, dir < 0 ? c->prog_fft_fwd[sys] : c->prog_fft_inv[sys], 0);
  if (rc) return rc;
  return copy_out(c, inout, c->scratch[sys], b);
}

int scft_calc_poly_density(scft_ctx* c) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  rc = stage_poly_density(c);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->st));
  return SCFT_OK;
}

int scft_update_1s(scft_ctx* c) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  rc = stage_update(c, 1);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->st));
  return SCFT_OK;
}

int scft_update_euler(scft_ctx* c) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  if (c->any_pairs) return fail(SCFT_EUNSUPPORTED, "the Euler update has no BC/AC field pairs");
  rc = stage_update(c, 0);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->st));
  return SCFT_OK;
}

int scft_calc_H(scft_ctx* c, double* H) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  if (!H) return fail(SCFT_EINVAL, "null output");
  const int ns = (int)c->sys.size();
  HIPCK(scft::launch_reduce(scft::RED_H, 0, c->d_sys, ns, c->nblk, c->st));
  int nan = 0;
  for (int i = 0; i < ns; i++) {
    SysHost& s = c->sys[i];
    double2 red[scft::RED_NSLOTS];
    if ((rc = copy_out(c, red, s.dev.red, sizeof red))) return rc;
    const cd field = cd(red[scft::RED_H].x, red[scft::RED_H].y) * s.dV;
    const cd Qs = cd(red[scft::RED_Q].x, red[scft::RED_Q].y) * (s.dV / s.info.V);
    cd logQ = std::log(Qs) - double(s.N) * red[scft::RED_SHIFT].y;
    cd h = field;
    if (s.info.nD > 0.0) h += -s.info.nD * logQ;
    // chains grafted to the plates, code/calc_h.cpp:31-35
    if (s.grafted) h += -double(s.p.n_exp_nr) * s.info.ng_per_np * qga(s, red).real();
    H[2 * i] = h.real();
    H[2 * i + 1] = h.imag();
    if (std::isnan(h.real()) || std::isnan(h.imag())) nan = 1;
  }
  if (nan) return fail(SCFT_ENAN, "H is NaN");
  return SCFT_OK;
}

int scft_synchronize(scft_ctx* c) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  HIPCK(hipDeviceSynchronize());
  return SCFT_OK;
}

int scft_time_kernel(scft_ctx* c, int which, int reps, double* avg_ms, double* bytes) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  if (reps < 1 || !avg_ms || !bytes) return fail(SCFT_EINVAL, "bad arguments");
  const int ns = (int)c->sys.size();
  double b = 0.0;
  if (c->ckpt && which != 3) return fail(SCFT_EINVAL, "with checkpointed propagators only which = 3 is timed");
  auto run_once = [&]() -> int {
    if (which == 0 || which == 1) {
      // the launch the iteration runs: one chain when the chains are split
      const Program& p = c->split_chains ? c->prog_prop_q : c->prog_prop;
      const size_t k = which == 0 ? 0 : 1;
      if (k >= p.launches.size()) return fail(SCFT_EINVAL, "this grid has no separate z pass");
      HIPCK(launch_one(c, p.launches[k], 1));
      return SCFT_OK;
    }
    if (which == 2) {
      HIPCK(scft::launch_density(c->prec, c->d_sys, ns, c->maxM, c->st));
      return SCFT_OK;
    }
    if (which == 4) {
      const SysHost& s0 = c->sys[0];
      const long n16 = (long)(2 * (size_t)s0.M * c->esz / 16);
      HIPCK(scft::launch_copy16(s0.dev.q, (char*)s0.dev.q + (size_t)2 * s0.M * c->esz, n16, c->st));
      return SCFT_OK;
    }
    return stage_update(c, c->sys[0].p.update_scheme);
  };
  if (which == 4 && c->sys[0].N < 4) return fail(SCFT_EINVAL, "the copy needs N >= 4");
  // algorithmic bytes of one launch
  if (which == 4) b = 4.0 * c->sys[0].M * (double)c->esz;
  for (auto& s : c->sys) {
    if (which == 4) break;
    const double M = s.M, pe = (double)c->esz;
    const int chains = (s.dev.qd != s.dev.q && !c->split_chains) ? 2 : 1;
    const bool two_pass = s.nz > 1;
    if (which == 0) b += chains * M * pe * (two_pass ? 2.0 : 3.0);
    else if (which == 1) b += chains * M * pe * 3.0;
    else if (which == 2) b += M * (2.0 * s.N * pe + 2 * 16.0 + 2 * 16.0);
    else {
      const double N = s.N;
      b += M * (10.0 * (N - 1) * pe + 2 * pe + 2 * N * pe + 2 * 16.0 + 24 * 16.0 + 12 * 16.0);
    }
  }
  if ((rc = run_once())) return rc;  // warm
  hipEvent_t e0, e1;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  HIPCK(hipEventRecord(e0, c->st));
  for (int r = 0; r < reps && rc == SCFT_OK; r++) rc = run_once();
  HIPCK(hipEventRecord(e1, c->st));
  HIPCK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *avg_ms = ms / reps;
  *bytes = b;
  return SCFT_OK;
}

int scft_describe(scft_ctx* c, char* buf, int len) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  if (!buf || len < 1) return fail(SCFT_EINVAL, "bad buffer");
  std::string o;
  auto prog = [&](const char* name, const Program& p) {
    o += name;
    o += ":";
    for (const Launch& L : p.launches) {
      char tmp[256];
      if (L.fast == 1)
        snprintf(tmp, sizeof tmp, " plane_conv[%dx%d prec%d jobs%d variant %d of %zu]", L.n0, L.n1, L.prec,
                 L.njobs, L.variants[L.choice], L.variants.size());
      else if (L.fast == 2)
        snprintf(tmp, sizeof tmp, " z_conv[%d tile%d prec%d jobs%d variant %d of %zu]", L.n0, L.tiles[L.choice],
                 L.prec, L.njobs, L.variants[L.choice], L.variants.size());
      else if (L.fast == 3)
        snprintf(tmp, sizeof tmp, " odd_line[%d axis%d lines%d mode%d prec%d jobs%d variant %d of %zu]", L.n0,
                 L.n1, L.tiles[L.choice], L.mode, L.prec, L.njobs, L.variants[L.choice], L.variants.size());
      else
        snprintf(tmp, sizeof tmp, " %s[mode%d prec%d jobs%d tb%d lds%zu]", L.plane ? "plane" : "line", L.mode,
                 L.prec, L.njobs, L.tb, L.lds);
      o += tmp;
      if (!L.tuned_ms.empty()) {
        o += " tuned_us=";
        for (float t : L.tuned_ms) {
          snprintf(tmp, sizeof tmp, "%.1f,", 1000.0 * t);
          o += tmp;
        }
      }
    }
    o += "\n";
  };
  prog("smear", c->prog_smear);
  if (c->ckpt) {
    prog("prop_checkpointed_q", c->prog_ck_q);
    prog("prop_checkpointed_qdag", c->prog_ck_qd);
  } else if (c->split_chains) {
    prog("prop_q", c->prog_prop_q);
    prog("prop_qdag", c->prog_prop_qd);
  } else {
    prog("prop", c->prog_prop);
  }
  prog("field_fwd", c->prog_fwd);
  prog("field_inv", c->prog_inv);
  {
    // propagator storage of system 0
    const SysHost& s0 = c->sys[0];
    const double slice = (double)s0.M * c->esz;
    const int slices = c->ckpt ? 2 * (c->ck_K + 1) + c->ck_ne : (s0.dev.qd != s0.dev.q ? 2 : 1) * s0.N;
    char tmp[256];
    if (c->ckpt)
      snprintf(tmp, sizeof tmp, "propagator_storage: %d slices, %.1f MiB (checkpointed, K=%d, %d segments)\n",
               slices, slices * slice / 1048576.0, c->ck_K, c->ck_ne);
    else
      snprintf(tmp, sizeof tmp, "propagator_storage: %d slices, %.1f MiB\n", slices, slices * slice / 1048576.0);
    o += tmp;
    snprintf(tmp, sizeof tmp, "device_bytes: %zu\n", c->dev_bytes);
    o += tmp;
  }
  snprintf(buf, len, "%s", o.c_str());
  return SCFT_OK;
}

int scft_iterate(scft_ctx* c, int n_iter) {
  int rc = check_ctx(c, 0);
  if (rc) return rc;
  if (n_iter < 0) return fail(SCFT_EINVAL, "negative iteration count");
  const int scheme = c->sys[0].p.update_scheme;
  if (!c->graph || c->graph_scheme != scheme) {
    // the first iteration runs eagerly (sets kernel attributes), the
    // second is captured
    if (n_iter == 0) return SCFT_OK;
    if ((rc = stage_iteration(c, scheme))) return rc;
    n_iter--;
    if (c->graph) {
      (void)hipGraphExecDestroy(c->graph);
      c->graph = nullptr;
    }
    hipGraph_t g;
    HIPCK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    rc = stage_iteration(c, scheme);
    hipError_t e = hipStreamEndCapture(c->st, &g);
    if (rc) {
      if (e == hipSuccess) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(SCFT_EDEVICE, "graph capture failed");
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(SCFT_EDEVICE, "graph instantiation failed");
    c->graph_scheme = scheme;
  }
  for (int it = 0; it < n_iter; it++) HIPCK(hipGraphLaunch(c->graph, c->st));
  HIPCK(hipStreamSynchronize(c->st));
  return SCFT_OK;
}

}  // extern "C"
