// Line passes for the odd lengths of the published 2D grids (525 x 875,
// 625 x 675; code/read_input.cpp:114-118 reads them from bcp.input).
//
// Every such length is n = N1 N2 with N1, N2 <= 35 products of 3, 5 and 7,
// and the line transform is split in two register-resident steps with one
// LDS exchange (Bailey's four-step ordering without the transpose):
//   j = N2 j1 + j2, k = k1 + N1 k2,
//   X[k1 + N1 k2] = sum_j2 W_N2^(j2 k2) W_n^(j2 k1) sum_j1 x[N2 j1 + j2] W_N1^(j1 k1)
// Step A: lane j2 of a line holds x[N2 j1 + j2] (j1 < N1), does the N1-point
// DFT in registers and the twiddle W_n^(j2 k1), and writes row k1 of the
// line tile in LDS.  Step B: lane k1 reads row k1 (N2 values) and does the
// N2-point DFT, which leaves X[k1 + N1 k2] in natural order in its
// registers.  The convolution of the propagator and smearing multiplies by
// the axis kernel there and runs the mirrored inverse (step B', LDS, step
// A'), so one pass is two exchanges and no extra global traffic.
//
// Loads and stores are coalesced in both layouts: along x (contiguous lines)
// lanes j2 read consecutive points; along y (stride nx) a block owns CPB
// adjacent columns and lanes of one point index read CPB consecutive values.
// The small DFTs are fully unrolled mixed-radix transforms with compile-time
// twiddles (cos / sin from a Taylor series on [0, pi/2) after an exact
// quadrant reduction in integers).
#include "kernels.h"
#include "lds_fft.hpp"
#include "scft_dev.hpp"

namespace scft {
namespace {

constexpr double kHalfPi = 1.57079632679489661923132169163975144;

// Taylor series on [0, pi/2): the terms fall from the first one on
constexpr double t_sin(double x) {
  double x2 = x * x, t = x, s = x;
  for (int k = 1; k <= 12; k++) {
    t *= -x2 / double((2 * k) * (2 * k + 1));
    s += t;
  }
  return s;
}
constexpr double t_cos(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int k = 1; k <= 12; k++) {
    t *= -x2 / double((2 * k - 1) * (2 * k));
    s += t;
  }
  return s;
}

// cos and sin of 2 pi m / N, m < N: 4 m = q N + r, angle = (q + r / N) pi / 2
template <int N>
struct Cis {
  double c[N], s[N];
  constexpr Cis() : c(), s() {
    for (int m = 0; m < N; m++) {
      const int q = (4 * m) / N, r = 4 * m - q * N;
      const double x = kHalfPi * double(r) / double(N);
      const double cx = t_cos(x), sx = t_sin(x);
      switch (q & 3) {
        case 0: c[m] = cx; s[m] = sx; break;
        case 1: c[m] = -sx; s[m] = cx; break;
        case 2: c[m] = -cx; s[m] = -sx; break;
        default: c[m] = sx; s[m] = -cx; break;
      }
    }
  }
};
template <int N>
struct CisTab {
  static constexpr Cis<N> v = Cis<N>();
};

// a[k] <- sum_j a[j] exp(SIGN 2 pi i j k / N) in registers: leaves 2..8 from
// lds_fft.hpp, composites by one Cooley-Tukey split N = P Q
// (j = Q j1 + j2, k = k1 + P k2) with compile-time twiddles.
template <int N, int SIGN, typename R>
__device__ __forceinline__ void reg_dft(cpx<R>* a) {
  if constexpr (N == 1) {
  } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 7 || N == 8) {
    dft<N, SIGN>(a);
  } else {
    constexpr int P = (N % 5 == 0) ? 5 : (N % 3 == 0) ? 3 : (N % 7 == 0) ? 7 : (N % 4 == 0) ? 4 : 2;
    constexpr int Q = N / P;
    cpx<R> t[N];
#pragma unroll
    for (int j2 = 0; j2 < Q; j2++) {
      cpx<R> u[P];
#pragma unroll
      for (int j1 = 0; j1 < P; j1++) u[j1] = a[Q * j1 + j2];
      dft<P, SIGN>(u);
#pragma unroll
      for (int k1 = 0; k1 < P; k1++) {
        const int m = j2 * k1;  // < N
        if (m == 0) t[j2 * P + k1] = u[k1];
        else
          t[j2 * P + k1] = cmul(u[k1], mk(R(CisTab<N>::v.c[m]), R(SIGN * CisTab<N>::v.s[m])));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) {
      cpx<R> v[Q];
#pragma unroll
      for (int j2 = 0; j2 < Q; j2++) v[j2] = t[j2 * P + k1];
      reg_dft<Q, SIGN>(v);
#pragma unroll
      for (int k2 = 0; k2 < Q; k2++) a[k1 + P * k2] = v[k2];
    }
  }
}

// Column-group index of a block of a strided (y) pass.  Workgroups are
// dispatched to the 8 XCDs in turn by linear id, and each XCD has its own
// L2.  Neighbouring column groups share 128-byte lines, so the blocks of one
// XCD get one contiguous run of groups: XCD x takes groups [x q + min(x, r),
// ...) with nb = 8 q + r.  A bijection of [0, nb); the launcher pads nb to a
// multiple of 8 so that block b of every job y sits on XCD b mod 8.
__device__ __forceinline__ int xcd_group(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

// LDS bound of a two-step block (gfx950: 160 KiB per workgroup)
constexpr size_t kOddLdsMax = 160 * 1024;

template <int N1, int N2, int AXIS, int CPB, int LW>
struct OddCfg {
  static constexpr int n = N1 * N2;
  static constexpr int J = N1 > N2 ? N1 : N2;  // lanes one line needs
  static constexpr int T = AXIS == 0 ? CPB * LW : ((CPB * J + 63) / 64) * 64;
};

template <typename R>
__device__ __forceinline__ cpx<R> conj_tw(cpx<R> t) {
  return mk(t.x, -t.y);
}

// One pass along AXIS (0: x lines, contiguous; 1: y lines, stride nx) over
// CPB lines per block.  MODE_CONV: forward, times g (the axis kernel, natural
// frequency order), inverse; MODE_FWD / MODE_INV: one transform, natural
// order in and out.  The last pass of a job applies post (exp(-w)) and scale.
template <typename R, int MODE, int N1, int N2, int AXIS, int CPB, int LW>
__global__ __launch_bounds__((OddCfg<N1, N2, AXIS, CPB, LW>::T)) void k_odd_line(const PassDesc* __restrict__ descs,
                                                                                int step) {
  using C = OddCfg<N1, N2, AXIS, CPB, LW>;
  constexpr int n = C::n;
  // CPB lines of n values, dynamic: up to 160 KiB per workgroup on gfx950
  extern __shared__ __align__(16) unsigned char odd_smem[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(odd_smem);
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  const long M = (long)d.nx * d.ny * d.nz;
  const int nlines = (int)(M / n);
  const int bx = AXIS == 0 ? (int)blockIdx.x : xcd_group(blockIdx.x, gridDim.x);
  if (bx * CPB >= nlines) return;
  const int t = threadIdx.x;
  int c, j;
  if constexpr (AXIS == 0) {
    c = t / LW;
    j = t - c * LW;
  } else {
    c = t % CPB;
    j = t / CPB;
  }
  const int line = bx * CPB + c;
  const bool live = line < nlines;
  long base, S;
  if constexpr (AXIS == 0) {
    base = (long)line * n;
    S = 1;
  } else {
    base = (long)(line % d.nx) + (long)d.nx * d.ny * (line / d.nx);
    S = d.nx;
  }
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step;
  const cpx<R>* tw = reinterpret_cast<const cpx<R>*>(d.tw0);
  const R* g = reinterpret_cast<const R*>(d.g0);
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp);
    scale = R(d.scale);
  }
  cpx<R>* L = lds + c * n;  // L[k1 * N2 + j2]
  auto finish = [&](long gi, cpx<R> o) {
    if (post) o = cmul(o, post[gi]);
    if (scale != R(1)) o = cscale(o, scale);
    out[gi] = o;
  };

  if constexpr (MODE == MODE_INV) {
    // step B': lane k1 = j holds X[k1 + N1 k2]
    if (live && j < N1) {
      cpx<R> b[N2];
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) b[k2] = in[base + (long)(j + N1 * k2) * S];
      reg_dft<N2, +1>(b);
#pragma unroll
      for (int j2 = 1; j2 < N2; j2++) b[j2] = cmul(b[j2], conj_tw(tw[j2 * j]));
#pragma unroll
      for (int j2 = 0; j2 < N2; j2++) L[j * N2 + j2] = b[j2];
    }
    __syncthreads();
    // step A': lane j2 = j
    if (live && j < N2) {
      cpx<R> a[N1];
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) a[k1] = L[k1 * N2 + j];
      reg_dft<N1, +1>(a);
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++) finish(base + (long)(N2 * j1 + j) * S, a[j1]);
    }
  } else {
    // step A: lane j2 = j holds x[N2 j1 + j2]
    if (live && j < N2) {
      cpx<R> a[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++) a[j1] = in[base + (long)(N2 * j1 + j) * S];
      reg_dft<N1, -1>(a);
#pragma unroll
      for (int k1 = 1; k1 < N1; k1++) a[k1] = cmul(a[k1], tw[j * k1]);
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) L[k1 * N2 + j] = a[k1];
    }
    __syncthreads();
    // step B: lane k1 = j
    if (live && j < N1) {
      cpx<R> b[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; j2++) b[j2] = L[j * N2 + j2];
      reg_dft<N2, -1>(b);
      if constexpr (MODE == MODE_FWD) {
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) finish(base + (long)(j + N1 * k2) * S, b[k2]);
      } else {
        if (g) {
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) b[k2] = cscale(b[k2], g[j + N1 * k2]);
        }
        reg_dft<N2, +1>(b);
#pragma unroll
        for (int j2 = 1; j2 < N2; j2++) b[j2] = cmul(b[j2], conj_tw(tw[j2 * j]));
        // lane k1 rewrites only its own row: no barrier before the store
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++) L[j * N2 + j2] = b[j2];
      }
    }
    if constexpr (MODE == MODE_CONV) {
      __syncthreads();
      // step A'
      if (live && j < N2) {
        cpx<R> a[N1];
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) a[k1] = L[k1 * N2 + j];
        reg_dft<N1, +1>(a);
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) finish(base + (long)(N2 * j1 + j) * S, a[j1]);
      }
    }
  }
}

template <typename R, int MODE, int N1, int N2, int AXIS, int CPB, int LW>
hipError_t launch_odd_m(const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  using C = OddCfg<N1, N2, AXIS, CPB, LW>;
  constexpr size_t lds = sizeof(cpx<R>) * CPB * C::n;
  static bool attr[64] = {};
  raise_lds_limit((const void*)k_odd_line<R, MODE, N1, N2, AXIS, CPB, LW>, lds, attr);
  hipLaunchKernelGGL((k_odd_line<R, MODE, N1, N2, AXIS, CPB, LW>), dim3(gridx, njobs), dim3(C::T), lds, st, d,
                     step);
  return hipGetLastError();
}

template <typename R, int N1, int N2, int AXIS, int CPB, int LW>
hipError_t launch_odd_t(int mode, const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  if (mode == MODE_CONV) return launch_odd_m<R, MODE_CONV, N1, N2, AXIS, CPB, LW>(d, njobs, gridx, step, st);
  if (mode == MODE_FWD) return launch_odd_m<R, MODE_FWD, N1, N2, AXIS, CPB, LW>(d, njobs, gridx, step, st);
  return launch_odd_m<R, MODE_INV, N1, N2, AXIS, CPB, LW>(d, njobs, gridx, step, st);
}

// ---------------------------------------------------------------------------
// Stockham variant: the line transform as a sequence of radix-R passes in
// LDS with compile-time radices, one butterfly per lane per pass (lanes per
// line = max n / R), so each lane does a few small DFTs instead of two large
// ones.  Pass p with radix R and NS = product of the earlier radices maps
// butterfly j to in x[j + r n/R], twiddle exp(sign 2 pi i r (j mod NS) /
// (NS R)), out y[(j / NS) NS R + (j mod NS) + r NS] (autosort Stockham, as
// lds_fft.hpp does with run-time radices).
// ---------------------------------------------------------------------------
template <int N>
struct RadixPlan {
  int np;
  int r[16];
  constexpr RadixPlan() : np(0), r() {
    int m = N;
    while (m > 1) {
      const int p = m % 5 == 0 ? 5 : m % 3 == 0 ? 3 : m % 7 == 0 ? 7 : m % 4 == 0 ? 4 : 2;
      r[np++] = p;
      m /= p;
    }
  }
  constexpr int max_bfly() const {
    int b = 0;
    for (int i = 0; i < np; i++) b = N / r[i] > b ? N / r[i] : b;
    return b;
  }
};
template <int N>
struct PlanTab {
  static constexpr RadixPlan<N> v = RadixPlan<N>();
};

template <int N, int AXIS, int CPB>
struct StockCfg {
  static constexpr int TPL = PlanTab<N>::v.max_bfly();  // lanes per line
  static constexpr int T = ((CPB * TPL + 63) / 64) * 64;
};

// passes P.. of one transform; g (if non-null) multiplies the last pass's
// outputs at their natural frequency index
template <typename R, int SIGN, int N, int P, int NS>
__device__ __forceinline__ void stock_passes(cpx<R>* __restrict__ L, const cpx<R>* __restrict__ tw,
                                             const R* __restrict__ g, int lane, bool on) {
  constexpr int NP = PlanTab<N>::v.np;
  if constexpr (P < NP) {
    constexpr int RAD = PlanTab<N>::v.r[P];
    constexpr int NB = N / RAD;
    constexpr int TWS = N / (NS * RAD);
    const bool act = on && lane < NB;
    cpx<R> a[RAD];
    const int j = lane, k = j % NS;
    if (act) {
#pragma unroll
      for (int r = 0; r < RAD; r++) a[r] = L[j + r * NB];
    }
    __syncthreads();
    if (act) {
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < RAD; r++) {
          cpx<R> t = tw[r * k * TWS];
          if (SIGN > 0) t.y = -t.y;
          a[r] = cmul(a[r], t);
        }
      }
      dft<RAD, SIGN>(a);
      const int base = (j / NS) * NS * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; r++) {
        const int i = base + r * NS;
        cpx<R> o = a[r];
        if (P + 1 == NP && g) o = cscale(o, g[i]);
        L[i] = o;
      }
    }
    __syncthreads();
    stock_passes<R, SIGN, N, P + 1, NS * RAD>(L, tw, g, lane, on);
  }
}

template <typename R, int MODE, int N, int AXIS, int CPB>
__global__ __launch_bounds__((StockCfg<N, AXIS, CPB>::T)) void k_stock_line(const PassDesc* __restrict__ descs,
                                                                           int step) {
  using C = StockCfg<N, AXIS, CPB>;
  constexpr int TPL = C::TPL;
  __shared__ cpx<R> lds[CPB * N];
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  const long M = (long)d.nx * d.ny * d.nz;
  const int nlines = (int)(M / N);
  const int bx = AXIS == 0 ? (int)blockIdx.x : xcd_group(blockIdx.x, gridDim.x);
  if (bx * CPB >= nlines) return;
  const int t = threadIdx.x;
  int c, lane;
  if constexpr (AXIS == 0) {
    c = t / TPL;
    lane = t - c * TPL;
  } else {
    c = t % CPB;
    lane = t / CPB;
  }
  const int line = bx * CPB + c;
  const bool on = c < CPB && lane < TPL && line < nlines;
  long base, S;
  if constexpr (AXIS == 0) {
    base = (long)line * N;
    S = 1;
  } else {
    base = (long)(line % d.nx) + (long)d.nx * d.ny * (line / d.nx);
    S = d.nx;
  }
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step;
  const cpx<R>* tw = reinterpret_cast<const cpx<R>*>(d.tw0);
  const R* g = MODE == MODE_CONV ? reinterpret_cast<const R*>(d.g0) : nullptr;
  cpx<R>* L = lds + (c < CPB ? c : 0) * N;
  if (on) {
    for (int i = lane; i < N; i += TPL) L[i] = in[base + (long)i * S];
  }
  __syncthreads();
  if constexpr (MODE == MODE_CONV) {
    stock_passes<R, -1, N, 0, 1>(L, tw, g, lane, on);
    stock_passes<R, +1, N, 0, 1>(L, tw, nullptr, lane, on);
  } else if constexpr (MODE == MODE_FWD) {
    stock_passes<R, -1, N, 0, 1>(L, tw, nullptr, lane, on);
  } else {
    stock_passes<R, +1, N, 0, 1>(L, tw, nullptr, lane, on);
  }
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp);
    scale = R(d.scale);
  }
  if (on) {
    for (int i = lane; i < N; i += TPL) {
      const long gi = base + (long)i * S;
      cpx<R> o = L[i];
      if (post) o = cmul(o, post[gi]);
      if (scale != R(1)) o = cscale(o, scale);
      out[gi] = o;
    }
  }
}

template <typename R, int MODE, int N, int AXIS, int CPB>
hipError_t launch_stock_m(const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  using C = StockCfg<N, AXIS, CPB>;
  hipLaunchKernelGGL((k_stock_line<R, MODE, N, AXIS, CPB>), dim3(gridx, njobs), dim3(C::T), 0, st, d, step);
  return hipGetLastError();
}

template <int NN, int AX, int CP>
hipError_t launch_stock_v(int prec, int mode, const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  if constexpr ((size_t)CP * NN * 16 <= 64 * 1024) {
    if (prec == 0) {
      if (mode == MODE_CONV) return launch_stock_m<double, MODE_CONV, NN, AX, CP>(d, njobs, gridx, step, st);
      if (mode == MODE_FWD) return launch_stock_m<double, MODE_FWD, NN, AX, CP>(d, njobs, gridx, step, st);
      return launch_stock_m<double, MODE_INV, NN, AX, CP>(d, njobs, gridx, step, st);
    }
  }
  if constexpr ((size_t)CP * NN * 8 <= 64 * 1024) {
    if (prec == 1) {
      if (mode == MODE_CONV) return launch_stock_m<float, MODE_CONV, NN, AX, CP>(d, njobs, gridx, step, st);
      if (mode == MODE_FWD) return launch_stock_m<float, MODE_FWD, NN, AX, CP>(d, njobs, gridx, step, st);
      return launch_stock_m<float, MODE_INV, NN, AX, CP>(d, njobs, gridx, step, st);
    }
  }
  return hipErrorInvalidValue;
}

// the precisions whose LDS (CPB lines of n values) fits a workgroup
template <int NN, int N1, int N2, int AX, int CP, int LW>
hipError_t launch_odd_v(int prec, int mode, const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  if constexpr ((size_t)CP * NN * 16 <= kOddLdsMax) {
    if (prec == 0) return launch_odd_t<double, N1, N2, AX, CP, LW>(mode, d, njobs, gridx, step, st);
  }
  if constexpr ((size_t)CP * NN * 8 <= kOddLdsMax) {
    if (prec == 1) return launch_odd_t<float, N1, N2, AX, CP, LW>(mode, d, njobs, gridx, step, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace

// X(n, N1, N2, axis, lines per block, lanes per line (axis 0))
#define SCFT_ODD_VARIANTS(X)                                                                         \
  X(625, 25, 25, 0, 2, 32) X(625, 25, 25, 0, 1, 64) X(625, 25, 25, 0, 2, 64)                         \
  X(525, 25, 21, 0, 2, 32) X(525, 25, 21, 0, 1, 64) X(525, 25, 21, 0, 2, 64)                         \
  X(675, 25, 27, 0, 2, 32) X(675, 25, 27, 0, 1, 64) X(675, 25, 27, 0, 2, 64)                         \
  X(875, 35, 25, 0, 1, 64) X(875, 35, 25, 0, 2, 64)                                                  \
  X(625, 25, 25, 1, 2, 0) X(625, 25, 25, 1, 4, 0) X(625, 25, 25, 1, 8, 0)                            \
  X(525, 25, 21, 1, 2, 0) X(525, 25, 21, 1, 4, 0) X(525, 25, 21, 1, 8, 0)                            \
  X(675, 25, 27, 1, 2, 0) X(675, 25, 27, 1, 4, 0) X(675, 25, 27, 1, 8, 0)                            \
  X(875, 35, 25, 1, 2, 0) X(875, 35, 25, 1, 4, 0) X(875, 35, 25, 1, 8, 0)

// S(n, axis, lines per block): Stockham variants (block size <= 1024)
#define SCFT_STOCK_VARIANTS(S)                                                                       \
  S(625, 0, 1) S(625, 0, 2) S(525, 0, 1) S(525, 0, 2) S(675, 0, 1) S(675, 0, 2) S(875, 0, 1)         \
  S(875, 0, 2) S(625, 1, 2) S(625, 1, 4) S(625, 1, 8) S(525, 1, 2) S(525, 1, 4) S(675, 1, 2)         \
  S(675, 1, 4) S(875, 1, 2) S(875, 1, 4)

int odd_line_variants(int prec, int n, int axis, int* ids, int* cpb, int max_ids) {
  int k = 0, id = 0;
  const size_t esz = prec == 0 ? 16 : 8;
  // LDS of one block: CPB lines of n complex values (two-step: dynamic, up
  // to 160 KiB; Stockham: static, up to 64 KiB)
#define X(nn, a, b, ax, cp, lw)                                                    \
  if (n == nn && axis == ax && k < max_ids && (size_t)cp * nn * esz <= kOddLdsMax) { \
    cpb[k] = cp;                                                                  \
    ids[k++] = id;                                                                \
  }                                                                               \
  id++;
  SCFT_ODD_VARIANTS(X)
#undef X
#define S(nn, ax, cp)                                                             \
  if (n == nn && axis == ax && k < max_ids && (size_t)cp * nn * esz <= 64 * 1024 && \
      StockCfg<nn, ax, cp>::T <= 1024) {                                          \
    cpb[k] = cp;                                                                  \
    ids[k++] = id;                                                                \
  }                                                                               \
  id++;
  SCFT_STOCK_VARIANTS(S)
#undef S
  return k;
}

hipError_t launch_odd_line(int prec, int mode, int variant, const PassDesc* d, int njobs, int gridx, int step,
                           hipStream_t st) {
  int id = 0;
#define X(nn, a, b, ax, cp, lw)                                                               \
  if (variant == id) return launch_odd_v<nn, a, b, ax, cp, lw>(prec, mode, d, njobs, gridx, step, st); \
  id++;
  SCFT_ODD_VARIANTS(X)
#undef X
#define S(nn, ax, cp)                                                                             \
  if (variant == id) {                                                                            \
    if constexpr (StockCfg<nn, ax, cp>::T <= 1024)                                                \
      return launch_stock_v<nn, ax, cp>(prec, mode, d, njobs, gridx, step, st);                   \
    return hipErrorInvalidValue;                                                                  \
  }                                                                                               \
  id++;
  SCFT_STOCK_VARIANTS(S)
#undef S
  return hipErrorInvalidValue;
}

}  // namespace scft
