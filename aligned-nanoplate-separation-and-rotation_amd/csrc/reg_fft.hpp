// Register-resident line transforms for gfx950.
//
// A line of n = G * E points is held by G consecutive lanes of one wavefront,
// lane t owning points x = t + G m (m = 0..E-1) in registers.  The forward
// DFT is done in four steps (n = G * E):
//   1. an E-point DFT of each lane's own registers (compile-time radices),
//   2. a per-lane twiddle exp(-2 pi i t k1 / n),
//   3. a G-point DFT across the lanes of the group: log2(G) radix-2
//      decimation-in-frequency stages, each one lane exchange, leaving
//      lane t with frequency k2 = bitrev(t),
// so lane t ends up holding X[k1 + E bitrev(t)] in register k1.  A pointwise
// multiplier over the spectrum is applied there and the inverse runs the
// steps backwards with conjugate twiddles; the result (times n) is back in
// the original layout.  Nothing touches LDS between the two transforms.
#pragma once

#include <hip/hip_runtime.h>

#include "lds_fft.hpp"

namespace scft {

// ---------------------------------------------------------------------------
// compile-time cos / sin of 2 pi j / n (octant reduction + Taylor series)
// ---------------------------------------------------------------------------
constexpr double kPiC = 3.14159265358979323846264338327950288;

constexpr double c_sin_series(double x) {
  double x2 = x * x, term = x, s = x;
  for (int i = 1; i < 16; i++) {
    term *= -x2 / double((2 * i) * (2 * i + 1));
    s += term;
  }
  return s;
}
constexpr double c_cos_series(double x) {
  double x2 = x * x, term = 1.0, s = 1.0;
  for (int i = 1; i < 16; i++) {
    term *= -x2 / double((2 * i - 1) * (2 * i));
    s += term;
  }
  return s;
}
// want_cos ? cos(2 pi j / n) : sin(2 pi j / n)
constexpr double c_trig(long j, long n, bool want_cos) {
  j %= n;
  if (j < 0) j += n;
  const long num = 8 * j;
  const long o = num / n;          // octant
  const long rem = num - o * n;    // angle = (pi/4) (o + rem / n)
  double s0 = 0.0, c0 = 1.0;       // sin, cos of the angle inside its quadrant
  if (o % 2 == 0) {
    if (rem != 0) {
      const double f = (kPiC / 4.0) * double(rem) / double(n);
      s0 = c_sin_series(f);
      c0 = c_cos_series(f);
    }
  } else {
    const double g = (kPiC / 4.0) * double(n - rem) / double(n);
    s0 = c_cos_series(g);
    c0 = c_sin_series(g);
  }
  const long q = (o / 2) % 4;
  double c = c0, s = s0;
  if (q == 1) {
    c = -s0;
    s = c0;
  } else if (q == 2) {
    c = -c0;
    s = -s0;
  } else if (q == 3) {
    c = s0;
    s = -c0;
  }
  return want_cos ? c : s;
}

template <int N>
struct TwTab {
  double c[N], s[N];
  constexpr TwTab() : c(), s() {
    for (int j = 0; j < N; j++) {
      c[j] = c_trig(j, N, true);
      s[j] = c_trig(j, N, false);
    }
  }
};

// multiply by exp(SIGN 2 pi i j / N), j a constant after unrolling
template <int N, int SIGN, typename R>
__device__ __forceinline__ cpx<R> twc(int j, cpx<R> v) {
  constexpr TwTab<N> tab{};
  const int jj = j % N;
  if (jj == 0) return v;
  const R c = R(tab.c[jj]);
  const R s = R(SIGN * tab.s[jj]);
  return mk(v.x * c - v.y * s, v.x * s + v.y * c);
}

// ---------------------------------------------------------------------------
// in-register DFT of compile-time length N (natural order in and out)
// ---------------------------------------------------------------------------
template <int N>
struct FirstRadix {
  static constexpr int value =
      (N % 4 == 0 && N != 8) ? 4
      : (N % 8 == 0)         ? 8
      : (N % 2 == 0)         ? 2
      : (N % 3 == 0)         ? 3
      : (N % 5 == 0)         ? 5
      : (N % 7 == 0)         ? 7
                             : N;
};

template <int N, int SIGN, typename R>
__device__ __forceinline__ void rdft(cpx<R>* a) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 7 || N == 8) {
    dft<N, SIGN>(a);
  } else {
    constexpr int P = FirstRadix<N>::value;
    static_assert(P == 2 || P == 3 || P == 4 || P == 5 || P == 7 || P == 8, "radix");
    constexpr int Q = N / P;
    cpx<R> b[P][Q];
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
      for (int q = 0; q < Q; q++) b[p][q] = a[p + P * q];
#pragma unroll
    for (int p = 0; p < P; p++) rdft<Q, SIGN>(b[p]);
#pragma unroll
    for (int k2 = 0; k2 < Q; k2++) {
      cpx<R> t[P];
#pragma unroll
      for (int p = 0; p < P; p++) t[p] = twc<N, SIGN>(p * k2, b[p][k2]);
      dft<P, SIGN>(t);
#pragma unroll
      for (int k1 = 0; k1 < P; k1++) a[k2 + Q * k1] = t[k1];
    }
  }
}

// ---------------------------------------------------------------------------
// lane exchange inside a group: the value of lane (lane ^ H), H <= 8, built
// from DPP moves that only permute inside quads, half rows and rows:
//   xor 1, 2, 3 : quad_perm;  xor 7 : row_half_mirror;  xor 15 : row_mirror;
//   xor 4 = (xor 7 then xor 3),  xor 8 = (xor 15 then xor 7).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int dpp_q1(int v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_q2(int v) { return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_q3(int v) { return __builtin_amdgcn_mov_dpp(v, 0x1B, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_hm(int v) { return __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_rm(int v) { return __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false); }

template <int H>
__device__ __forceinline__ int lane_xor_i(int v) {
  static_assert(H == 1 || H == 2 || H == 4 || H == 8, "lane exchange distance");
  if constexpr (H == 1) return dpp_q1(v);
  else if constexpr (H == 2) return dpp_q2(v);
  else if constexpr (H == 4) return dpp_q3(dpp_hm(v));
  else return dpp_hm(dpp_rm(v));
}
template <int H>
__device__ __forceinline__ float lane_xor(float v) {
  return __int_as_float(lane_xor_i<H>(__float_as_int(v)));
}
template <int H>
__device__ __forceinline__ double lane_xor(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_xor_i<H>((int)(b & 0xffffffffLL));
  const int hi = lane_xor_i<H>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int G>
__device__ __forceinline__ int bitrev_g(int t) {
  int r = 0;
#pragma unroll
  for (int b = 1; b < G; b <<= 1) {
    r = (r << 1) | (t & 1);
    t >>= 1;
  }
  return r;
}

// forward cross-lane stage with half size H: lo lanes get own + partner,
// hi lanes (partner - own) * exp(-2 pi i j / 2H), j = t mod H
template <typename R, int NL, int E, int H>
__device__ __forceinline__ void cross_fwd_stage(cpx<R>* v, int t, const cpx<R>* __restrict__ tw) {
  const bool hi = (t & H) != 0;
  const R sgn = hi ? R(-1) : R(1);
  cpx<R> w = mk(R(1), R(0));
  if (hi && H > 1) w = tw[(t & (H - 1)) * (NL / (2 * H))];
#pragma unroll
  for (int k = 0; k < E; k++) {
    const cpx<R> own = v[k];
    const cpx<R> part = mk(lane_xor<H>(own.x), lane_xor<H>(own.y));
    cpx<R> r = mk(part.x + sgn * own.x, part.y + sgn * own.y);
    if (H > 1) r = cmul(r, w);
    v[k] = r;
  }
}

// inverse of cross_fwd_stage (times 2): hi lanes first multiply by the
// conjugate twiddle, then lo: own + partner, hi: partner - own
template <typename R, int NL, int E, int H>
__device__ __forceinline__ void cross_inv_stage(cpx<R>* v, int t, const cpx<R>* __restrict__ tw) {
  const bool hi = (t & H) != 0;
  const R sgn = hi ? R(-1) : R(1);
  cpx<R> w = mk(R(1), R(0));
  if (hi && H > 1) {
    w = tw[(t & (H - 1)) * (NL / (2 * H))];
    w.y = -w.y;
  }
#pragma unroll
  for (int k = 0; k < E; k++) {
    cpx<R> x = v[k];
    if (H > 1) x = cmul(x, w);
    const cpx<R> part = mk(lane_xor<H>(x.x), lane_xor<H>(x.y));
    v[k] = mk(part.x + sgn * x.x, part.y + sgn * x.y);
  }
}

template <typename R, int NL, int E, int G>
__device__ __forceinline__ void cross_fwd(cpx<R>* v, int t, const cpx<R>* tw) {
  if constexpr (G >= 16) cross_fwd_stage<R, NL, E, 8>(v, t, tw);
  if constexpr (G >= 8) cross_fwd_stage<R, NL, E, 4>(v, t, tw);
  if constexpr (G >= 4) cross_fwd_stage<R, NL, E, 2>(v, t, tw);
  if constexpr (G >= 2) cross_fwd_stage<R, NL, E, 1>(v, t, tw);
}
template <typename R, int NL, int E, int G>
__device__ __forceinline__ void cross_inv(cpx<R>* v, int t, const cpx<R>* tw) {
  if constexpr (G >= 2) cross_inv_stage<R, NL, E, 1>(v, t, tw);
  if constexpr (G >= 4) cross_inv_stage<R, NL, E, 2>(v, t, tw);
  if constexpr (G >= 8) cross_inv_stage<R, NL, E, 4>(v, t, tw);
  if constexpr (G >= 16) cross_inv_stage<R, NL, E, 8>(v, t, tw);
}

// Forward DFT of one line held by G lanes: on return lane t holds
// X[k + E bitrev(t)] in register k (unnormalised).
template <typename R, int NL, int G>
__device__ __forceinline__ void group_fwd(cpx<R>* v, int t, const cpx<R>* __restrict__ tw) {
  constexpr int E = NL / G;
  rdft<E, -1>(v);
  if constexpr (G > 1) {
#pragma unroll
    for (int k = 1; k < E; k++) v[k] = cmul(v[k], tw[t * k]);
    cross_fwd<R, NL, E, G>(v, t, tw);
  }
}

// Inverse of group_fwd (unnormalised backward DFT): takes lane t holding
// X[k + E bitrev(t)] in register k, returns the points x = t + G m.
template <typename R, int NL, int G>
__device__ __forceinline__ void group_inv(cpx<R>* v, int t, const cpx<R>* __restrict__ tw) {
  constexpr int E = NL / G;
  if constexpr (G > 1) {
    cross_inv<R, NL, E, G>(v, t, tw);
#pragma unroll
    for (int k = 1; k < E; k++) {
      cpx<R> w = tw[t * k];
      w.y = -w.y;
      v[k] = cmul(v[k], w);
    }
  }
  rdft<E, +1>(v);
}

// Convolution of one line held by G lanes: IDFT[g(k) DFT[a]], g holding
// 1/n.  `tw` is exp(-2 pi i m / n), m = 0..n-1, and `g` the multiplier over k
// (both usually staged in LDS).
template <typename R, int NL, int G>
__device__ __forceinline__ void group_conv(cpx<R>* v, int t, const cpx<R>* __restrict__ tw,
                                           const R* __restrict__ g) {
  static_assert(NL % G == 0, "lanes must divide the line");
  constexpr int E = NL / G;
  group_fwd<R, NL, G>(v, t, tw);
  const int kb = E * bitrev_g<G>(t);
#pragma unroll
  for (int k = 0; k < E; k++) v[k] = cscale(v[k], g[kb + k]);
  group_inv<R, NL, G>(v, t, tw);
}

// One line transform in the lane-group layout: MODE 0 convolution,
// 1 forward DFT, 2 backward DFT.  Point b + o of the line is read with
// ld(b, o) and written with st(b, o, value), b the lane's first index (t,
// or E bitrev(t) in the frequency layout) and o an offset known after
// unrolling, so accessors can put the lane part in one register and fold
// the rest into immediate or scalar offsets.  Forward output and backward
// input are in natural frequency order.
template <int MODE, typename R, int NL, int G, typename Ld, typename St>
__device__ __forceinline__ void line_op_io(int t, const cpx<R>* __restrict__ tw, const R* __restrict__ g,
                                           Ld ld, St st) {
  constexpr int E = NL / G;
  cpx<R> v[E];
  if (MODE == 2) {
    const int kb = E * bitrev_g<G>(t);
#pragma unroll
    for (int k = 0; k < E; k++) v[k] = ld(kb, k);
    group_inv<R, NL, G>(v, t, tw);
#pragma unroll
    for (int m = 0; m < E; m++) st(t, G * m, v[m]);
  } else {
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = ld(t, G * m);
    if (MODE == 0) {
      group_conv<R, NL, G>(v, t, tw, g);
#pragma unroll
      for (int m = 0; m < E; m++) st(t, G * m, v[m]);
    } else {
      group_fwd<R, NL, G>(v, t, tw);
      const int kb = E * bitrev_g<G>(t);
#pragma unroll
      for (int k = 0; k < E; k++) st(kb, k, v[k]);
    }
  }
}

// line_op_io reading and writing in place through `at(i)`, the address of
// point i of the line.
template <int MODE, typename R, int NL, int G, typename At>
__device__ __forceinline__ void line_op(int t, const cpx<R>* __restrict__ tw, const R* __restrict__ g, At at) {
  line_op_io<MODE, R, NL, G>(
      t, tw, g, [&](int b, int o) { return *at(b + o); }, [&](int b, int o, cpx<R> v) { *at(b + o) = v; });
}

// ---------------------------------------------------------------------------
// Raw buffer access (MUBUF) to one array of complex values: the base is a
// wave-uniform resource, the lane part of the address one 32-bit register
// (voff) and the rest a scalar offset (soff), all in bytes.  Accesses past
// `bytes` read zero and drop writes.
// ---------------------------------------------------------------------------
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, unsigned bytes) {
  // word 3 = 0x00020000: the raw-buffer format of the gfx9 family
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename R>
__device__ __forceinline__ cpx<R> buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ cpx<float> buf_ld<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const u32x2_t u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return mk(__uint_as_float(u.x), __uint_as_float(u.y));
}
template <>
__device__ __forceinline__ cpx<double> buf_ld<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const u32x4_t u = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return mk(__hiloint2double((int)u.y, (int)u.x), __hiloint2double((int)u.w, (int)u.z));
}

__device__ __forceinline__ void buf_st(cpx<float> v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  u32x2_t u;
  u.x = __float_as_uint(v.x);
  u.y = __float_as_uint(v.y);
  __builtin_amdgcn_raw_buffer_store_b64(u, r, voff, soff, 0);
}
__device__ __forceinline__ void buf_st(cpx<double> v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  u32x4_t u;
  u.x = (unsigned)__double2loint(v.x);
  u.y = (unsigned)__double2hiint(v.x);
  u.z = (unsigned)__double2loint(v.y);
  u.w = (unsigned)__double2hiint(v.y);
  __builtin_amdgcn_raw_buffer_store_b128(u, r, voff, soff, 0);
}

}  // namespace scft
