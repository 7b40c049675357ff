// In-LDS mixed-radix Stockham FFT engine for gfx950 (64-wide wavefronts).
//
// A workgroup owns a tile of B lines of length n that sits in LDS.  Element
// (line l, point i) lives at lds[l * ls + i * ps]: ls == 1 is the
// point-major layout used for strided axes (consecutive lanes touch
// consecutive lines), ps == 1 the line-major layout used along x.
//
// Each radix pass is done in place with register staging: every thread
// gathers the inputs of all butterflies it owns, the workgroup
// synchronises, then each thread writes its outputs to their Stockham
// positions.  One LDS buffer per tile therefore suffices, which keeps a whole
// 64x64 complex128 plane (64 KiB) resident with room for a second workgroup.
//
// Pass p with radix R and Ns = product of the earlier radices maps butterfly
// j (0 <= j < n/R) of a line to
//     in : x[j + r n/R],                         r = 0..R-1
//     tw : exp(sign 2 pi i r (j mod Ns) / (Ns R))
//     out: y[(j / Ns) Ns R + (j mod Ns) + r Ns]
// (the autosort Stockham formulation of Govindaraju et al., SC'08), so the
// result of the last pass is in natural order.
#pragma once

#include <hip/hip_runtime.h>

namespace scft {

template <typename R>
struct cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> mk(R a, R b) {
  cpx<R> c;
  c.x = a;
  c.y = b;
  return c;
}
template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) {
  return mk(a.x - b.x, a.y - b.y);
}
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) {
  return mk(a.x * s, a.y * s);
}
// multiply by sign * i
template <int SIGN, typename R>
__device__ __forceinline__ cpx<R> mul_si(cpx<R> a) {
  return SIGN < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// ---------------------------------------------------------------------------
// Small DFTs: y[m] = sum_r a[r] exp(SIGN 2 pi i r m / R), in registers.
// ---------------------------------------------------------------------------
template <int SIGN, typename R>
__device__ __forceinline__ void dft2(cpx<R>* a) {
  cpx<R> t = a[0];
  a[0] = t + a[1];
  a[1] = t - a[1];
}

template <int SIGN, typename R>
__device__ __forceinline__ void dft4(cpx<R>* a) {
  cpx<R> t0 = a[0] + a[2], t1 = a[0] - a[2];
  cpx<R> t2 = a[1] + a[3], t3 = mul_si<SIGN>(a[1] - a[3]);
  a[0] = t0 + t2;
  a[2] = t0 - t2;
  a[1] = t1 + t3;
  a[3] = t1 - t3;
}

template <int SIGN, typename R>
__device__ __forceinline__ void dft8(cpx<R>* a) {
  const R h = R(0.70710678118654752440084436210484904);
  cpx<R> e[4] = {a[0], a[2], a[4], a[6]};
  cpx<R> o[4] = {a[1], a[3], a[5], a[7]};
  dft4<SIGN>(e);
  dft4<SIGN>(o);
  // W^m = exp(SIGN 2 pi i m / 8)
  cpx<R> w1 = cmul(o[1], mk(h, SIGN * h));
  cpx<R> w2 = mul_si<SIGN>(o[2]);
  cpx<R> w3 = cmul(o[3], mk(-h, SIGN * h));
  a[0] = e[0] + o[0];
  a[4] = e[0] - o[0];
  a[1] = e[1] + w1;
  a[5] = e[1] - w1;
  a[2] = e[2] + w2;
  a[6] = e[2] - w2;
  a[3] = e[3] + w3;
  a[7] = e[3] - w3;
}

// Odd prime radix P: with t_r = a_r + a_{P-r}, d_r = a_r - a_{P-r},
//   y_m     = a_0 + sum_r C[rm] t_r + SIGN i sum_r S[rm] d_r
//   y_{P-m} = a_0 + sum_r C[rm] t_r - SIGN i sum_r S[rm] d_r
// with C[j] = cos(2 pi j / P), S[j] = sin(2 pi j / P).
template <int P>
struct OddTab;
template <>
struct OddTab<3> {
  static constexpr double C[3] = {1.0, -0.5, -0.5};
  static constexpr double S[3] = {0.0, 0.86602540378443864676372317075294,
                                  -0.86602540378443864676372317075294};
};
template <>
struct OddTab<5> {
  static constexpr double C[5] = {1.0, 0.30901699437494742410229341718282,
                                  -0.80901699437494742410229341718282,
                                  -0.80901699437494742410229341718282,
                                  0.30901699437494742410229341718282};
  static constexpr double S[5] = {0.0, 0.95105651629515357211643933337938,
                                  0.58778525229247312916870595463907,
                                  -0.58778525229247312916870595463907,
                                  -0.95105651629515357211643933337938};
};
template <>
struct OddTab<7> {
  static constexpr double C[7] = {1.0,
                                  0.62348980185873353052500488400424,
                                  -0.22252093395631440428890256449679,
                                  -0.90096886790241912623610231950745,
                                  -0.90096886790241912623610231950745,
                                  -0.22252093395631440428890256449679,
                                  0.62348980185873353052500488400424};
  static constexpr double S[7] = {0.0,
                                  0.78183148246802980870844452667406,
                                  0.97492791218182360701813168299393,
                                  0.43388373911755812047576833284836,
                                  -0.43388373911755812047576833284836,
                                  -0.97492791218182360701813168299393,
                                  -0.78183148246802980870844452667406};
};

template <int P, int SIGN, typename R>
__device__ __forceinline__ void dft_odd(cpx<R>* a) {
  constexpr int H = (P - 1) / 2;
  cpx<R> t[H], d[H];
  cpx<R> y0 = a[0];
#pragma unroll
  for (int r = 1; r <= H; r++) {
    t[r - 1] = a[r] + a[P - r];
    d[r - 1] = a[r] - a[P - r];
    y0 = y0 + t[r - 1];
  }
  cpx<R> out[P];
  out[0] = y0;
#pragma unroll
  for (int m = 1; m <= H; m++) {
    cpx<R> re = a[0];
    cpx<R> im = mk(R(0), R(0));
#pragma unroll
    for (int r = 1; r <= H; r++) {
      const int j = (r * m) % P;
      re = re + cscale(t[r - 1], R(OddTab<P>::C[j]));
      im = im + cscale(d[r - 1], R(OddTab<P>::S[j]));
    }
    cpx<R> s = mul_si<SIGN>(im);
    out[m] = re + s;
    out[P - m] = re - s;
  }
#pragma unroll
  for (int m = 0; m < P; m++) a[m] = out[m];
}

template <int RAD, int SIGN, typename R>
__device__ __forceinline__ void dft(cpx<R>* a) {
  if constexpr (RAD == 2) dft2<SIGN>(a);
  else if constexpr (RAD == 4) dft4<SIGN>(a);
  else if constexpr (RAD == 8) dft8<SIGN>(a);
  else dft_odd<RAD, SIGN>(a);
}

// ---------------------------------------------------------------------------
// Radix plan of one line length: up to 8 passes, 4 bits per radix.
// ---------------------------------------------------------------------------
struct Plan {
  int n;
  int npass;
  unsigned radices;  // radix of pass p in bits [4p, 4p+4)
  __host__ __device__ int radix(int p) const { return (radices >> (4 * p)) & 15; }
};

// One in-place Stockham pass over a tile.  `tw` holds exp(-2 pi i m / n),
// m = 0..n-1 (conjugated on the fly for SIGN > 0).  If `g` is non-null the
// outputs are multiplied by g[i] (i = final frequency index) on the way out.
template <typename R, int RAD, int SIGN, int MAXV>
__device__ __forceinline__ void stockham_pass(cpx<R>* __restrict__ lds, int n, int B, int ls,
                                              int ps, int Ns, const cpx<R>* __restrict__ tw,
                                              const R* __restrict__ g, int tid, int T) {
  constexpr int UMAX = MAXV / RAD;
  const int nb = n / RAD;
  const int total = nb * B;
  const int twstep = n / (Ns * RAD);
  cpx<R> v[UMAX * RAD];
  int lpos[UMAX], jpos[UMAX];
#pragma unroll
  for (int u = 0; u < UMAX; u++) {
    const int w = tid + u * T;
    if (w < total) {
      int l, j;
      if (ls == 1) {
        l = w % B;
        j = w / B;
      } else {
        j = w % nb;
        l = w / nb;
      }
      lpos[u] = l;
      jpos[u] = j;
#pragma unroll
      for (int r = 0; r < RAD; r++) v[u * RAD + r] = lds[l * ls + (j + r * nb) * ps];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < UMAX; u++) {
    const int w = tid + u * T;
    if (w < total) {
      const int l = lpos[u], j = jpos[u];
      const int k = j % Ns;
      cpx<R>* a = v + u * RAD;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < RAD; r++) {
          cpx<R> t = tw[r * k * twstep];
          if (SIGN > 0) t.y = -t.y;
          a[r] = cmul(a[r], t);
        }
      }
      dft<RAD, SIGN>(a);
      const int base = (j / Ns) * Ns * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; r++) {
        const int i = base + r * Ns;
        cpx<R> o = a[r];
        if (g) o = cscale(o, g[i]);
        lds[l * ls + i * ps] = o;
      }
    }
  }
  __syncthreads();
}

template <typename R, int SIGN, int MAXV>
__device__ __forceinline__ void pass_dispatch(int rad, cpx<R>* lds, int n, int B, int ls, int ps,
                                              int Ns, const cpx<R>* tw, const R* g, int tid,
                                              int T) {
  switch (rad) {
    case 2: stockham_pass<R, 2, SIGN, MAXV>(lds, n, B, ls, ps, Ns, tw, g, tid, T); break;
    case 3: stockham_pass<R, 3, SIGN, MAXV>(lds, n, B, ls, ps, Ns, tw, g, tid, T); break;
    case 4: stockham_pass<R, 4, SIGN, MAXV>(lds, n, B, ls, ps, Ns, tw, g, tid, T); break;
    case 5: stockham_pass<R, 5, SIGN, MAXV>(lds, n, B, ls, ps, Ns, tw, g, tid, T); break;
    case 7: stockham_pass<R, 7, SIGN, MAXV>(lds, n, B, ls, ps, Ns, tw, g, tid, T); break;
    default: stockham_pass<R, 8, SIGN, MAXV>(lds, n, B, ls, ps, Ns, tw, g, tid, T); break;
  }
}

// Full 1D transform of every line of the tile.  For SIGN < 0 the optional
// multiplier g is applied on the last pass's outputs.
template <typename R, int SIGN, int MAXV>
__device__ __forceinline__ void lds_fft(const Plan& pl, cpx<R>* lds, int B, int ls, int ps,
                                        const cpx<R>* tw, const R* g, int tid, int T) {
  int Ns = 1;
  for (int p = 0; p < pl.npass; p++) {
    const int rad = pl.radix(p);
    pass_dispatch<R, SIGN, MAXV>(rad, lds, pl.n, B, ls, ps, Ns, tw,
                                 (p == pl.npass - 1) ? g : nullptr, tid, T);
    Ns *= rad;
  }
}

}  // namespace scft
