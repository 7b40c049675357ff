// Device-side descriptors shared by the kernels (scft_kernels.hip) and the
// host context (scft_api.hip).  All tables live in device memory and are
// built once per context; per-call variation (the contour step) is a scalar
// kernel argument.
#pragma once

#include <hip/hip_runtime.h>

#include "lds_fft.hpp"

namespace scft {

enum { MODE_CONV = 0, MODE_FWD = 1, MODE_INV = 2 };

// The `step` argument of a pass: the slice index that addresses the arrays
// (in = in + (slice + in_soff) * in_step) in the low 16 bits, and the
// contour index s that selects exp(-w(s)) and the jobs still active in the
// high 16 bits.  They are equal except in the checkpointed propagator, whose
// windows hold one segment of the chain.
__host__ __device__ constexpr int pack_step(int slice, int contour) { return slice | (contour << 16); }
__device__ __forceinline__ int step_slice(int step) { return step & 0xFFFF; }
__device__ __forceinline__ int step_contour(int step) { return step >> 16; }

// One array of one system for an FFT-style pass.
struct PassDesc {
  const void* in;
  void* out;
  long in_step, out_step;   // element offset per contour step
  int in_soff, out_soff;    // in = in + (step + in_soff) * in_step
  const void* post_a;       // optional pointwise multiplier on the last store
  const void* post_b;
  int post_switch;          // post = step < post_switch ? post_a : post_b
  int steps_active;         // job runs while step < steps_active (0: always)
  const void* g0;           // 1D multiplier of the pass axis (conv mode)
  const void* g1;           // plane pass: multiplier along y
  const void* tw0;          // twiddles of the pass axis, exp(-2 pi i m / n)
  const void* tw1;
  Plan p0, p1;
  int nx, ny, nz;
  int axis;                 // line pass axis: 0 x, 1 y, 2 z
  int tl;                   // lines per tile (line pass)
  int ntiles;
  int last;                 // this pass applies post and scale
  double scale;
};

// One plate: absolute centre and body axes (u normal, e1 e2 in-plane).
struct PlateDev {
  double c[3];
  double u[3], e1[3], e2[3];
};

// Per-system state for the pointwise kernels.
struct SysDev {
  int M, N, Na, nx, ny, nz;
  int dim, neg_ab, do_cl, np_chem;
  double nD, V, dV, rho0, kappa, chi_ab, lam_imag, lam_real;
  double fA;                // Na / N, weights of the gauge shift
  double2 *wpl, *wabp, *wabm;
  double2 *rhoA, *rhoB;
  double2 *wa, *wb;         // assembled fields, smeared in place -> smw
  double2 *rho_np;          // plate density, real space
  double2 *exp_nrH;         // its forward transform
  double *hhat_k;           // smearing kernel in k-space
  double *Xk;               // phiD rho0 N h^2 (gaa + 2 gab + gbb) (+ graft term)
  double *X2k;              // phiD rho0 N h^2 (gaa - gbb), negative-chi branch
  void *q, *qd;             // N * M propagators, propagator precision
                            // (checkpointed: slot 1 of the windows below)
  void *qlast;              // q_{N-1}
  // checkpointed propagators: windows of K + 1 slices of the q and q-dagger
  // chains and the q checkpoints (null otherwise)
  void *ck_sq, *ck_tq, *ck_cq;
  void *expA, *expB;        // exp(-(smw - shift)), propagator precision
  double2 *etap, *etam;     // noise (k-space): etap[0..M) w+, etap[M..2M) AB+, etam AB-
  unsigned long long *rng_ctr;  // iteration counter of the noise generator
  unsigned long long seed;
  // complex-Langevin density averages (code/array_utils.cpp:13-35): running
  // sums of rhoA, rhoB over the sampled iterations and their count
  double2 *avgA, *avgB;
  double *n_samples;
  int sample_wait, sample_freq;
  double noise_scale[3];    // sqrt(2 lambda) / prod sqrt(dx) of the three fields
  double2 *red;             // scalar slots, see RED_*
  double2 *partials;        // reduction scratch, nblk per slot
  int nblk;                 // partial blocks of a full-grid reduction
  // plates (code/initialize.cpp:578-612 and the square extension)
  int n_plates, np_type;
  double Lb[3], dxb[3];
  double len_nr, R_nr, xi_nr;
  PlateDev plate[2];
  double2 *grad_partials;   // 8 complex per block: plate derivatives of H
  // BC / AC field pairs (code/globals.h:76-90; null when the pair is off)
  // and the chains grafted to the plates (code/grafted_nanoparticle.cpp)
  double2 *wbcp, *wbcm, *wacp, *wacm;
  int neg_bc, neg_ac;
  double chi_bc, chi_ac;    // |chi|
  int grafted, Ng, n_exp;
  double ng_per_np;
  double2 *wg;              // assembled graft field, smeared in place -> smwg
  double2 *expl;            // graft sites, unit integral
  double2 *rhog;            // graft density rhog_exp
  void *qg, *qgd;           // Ng * M graft propagators, propagator precision
  void *expG;               // exp(-(smwg - c))
  double *XGk;              // phiG rho0 Ng h^2 gc
  double *XAk, *XBk;        // phiD rho0 N h^2 gaa, gbb
};

enum {
  RED_Q = 0,       // sum of q[N-1] (then Q = sum * dV / V, shifted gauge)
  RED_H = 1,       // sum of the field part of H
  RED_SHIFT = 2,   // sum of Re(fA smwa + (1 - fA) smwb); .y = shift c
  RED_RHO = 3,     // sum of rhoA + rhoB (segment count check)
  RED_VEXP = 4,    // sum of the plate density (volume bookkeeping)
  RED_QGA = 5,     // sum of expl_grafts log q_g[Ng-1] (shifted gauge)
  RED_NSLOTS = 6
};

}  // namespace scft
