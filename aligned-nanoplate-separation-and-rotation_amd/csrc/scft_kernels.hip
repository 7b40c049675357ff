// HIP kernels of the SCFT hot path for gfx950 and their launch wrappers.
//
// FFT-style passes (k_plane, k_line) run the in-LDS Stockham engine of
// lds_fft.hpp over tiles described by a device table of PassDesc, one table
// row per (array, system); blockIdx.y selects the row, blockIdx.x the tile.
// The propagator step q_{s+1} = exp(-w) IFFT[Phi(k) FFT q_s]
// (code/diblock_discrete.cpp:44-54) uses that Phi(k) = exp(-k^2/6) is a
// product of 1D Gaussians: the 3D convolution is done as 1D
// forward-FFT / multiply / inverse-FFT along x and y inside one LDS-resident
// xy plane, then along z in strided pencil tiles whose store applies exp(-w).
// Pointwise kernels take a device table of SysDev, blockIdx.y = system.
#include <hip/hip_runtime.h>

#include "../../include/scft.h"
#include "kernels.h"
#include "scft_dev.hpp"

namespace scft {

template <typename R>
__device__ __forceinline__ cpx<R> ld(const void* base, long i) {
  return reinterpret_cast<const cpx<R>*>(base)[i];
}

__device__ __forceinline__ double2 dadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 dscale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
__device__ __forceinline__ double2 ddiv(double2 a, double2 b) {
  // a / b with the scaling used by std::complex for well-ranged values
  const double den = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ __forceinline__ double2 dexp(double2 a) {
  double s, c;
  sincos(a.y, &s, &c);
  const double e = exp(a.x);
  return make_double2(e * c, e * s);
}
__device__ __forceinline__ double2 mul_i(double2 a) { return make_double2(-a.y, a.x); }

// ---------------------------------------------------------------------------
// Plane pass: x and y lines of one xy plane in LDS.
// ---------------------------------------------------------------------------
template <typename R, int MODE, int TB, int MAXV>
__global__ __launch_bounds__(TB) void k_plane(const PassDesc* __restrict__ descs, int step) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  const int z = blockIdx.x;
  if (z >= d.nz) return;
  const int tid = threadIdx.x;
  const int nx = d.nx, ny = d.ny, nxy = nx * ny;
  const long poff = (long)z * nxy;
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step + poff;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step + poff;
  for (int i = tid; i < nxy; i += TB) lds[i] = in[i];
  __syncthreads();
  const cpx<R>* tw0 = reinterpret_cast<const cpx<R>*>(d.tw0);
  const cpx<R>* tw1 = reinterpret_cast<const cpx<R>*>(d.tw1);
  const R* g0 = reinterpret_cast<const R*>(d.g0);
  const R* g1 = reinterpret_cast<const R*>(d.g1);
  // x lines: B = ny lines, line-major (ls = nx, ps = 1)
  if (MODE == MODE_CONV) {
    lds_fft<R, -1, MAXV>(d.p0, lds, ny, nx, 1, tw0, g0, tid, TB);
    lds_fft<R, +1, MAXV>(d.p0, lds, ny, nx, 1, tw0, nullptr, tid, TB);
  } else if (MODE == MODE_FWD) {
    lds_fft<R, -1, MAXV>(d.p0, lds, ny, nx, 1, tw0, nullptr, tid, TB);
  } else {
    lds_fft<R, +1, MAXV>(d.p0, lds, ny, nx, 1, tw0, nullptr, tid, TB);
  }
  // y lines: B = nx lines, point-major (ls = 1, ps = nx)
  if (ny > 1) {
    if (MODE == MODE_CONV) {
      lds_fft<R, -1, MAXV>(d.p1, lds, nx, 1, nx, tw1, g1, tid, TB);
      lds_fft<R, +1, MAXV>(d.p1, lds, nx, 1, nx, tw1, nullptr, tid, TB);
    } else if (MODE == MODE_FWD) {
      lds_fft<R, -1, MAXV>(d.p1, lds, nx, 1, nx, tw1, nullptr, tid, TB);
    } else {
      lds_fft<R, +1, MAXV>(d.p1, lds, nx, 1, nx, tw1, nullptr, tid, TB);
    }
  }
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp) + poff;
    scale = R(d.scale);
  }
  for (int i = tid; i < nxy; i += TB) {
    cpx<R> o = lds[i];
    if (post) o = cmul(o, post[i]);
    if (scale != R(1)) o = cscale(o, scale);
    out[i] = o;
  }
}

// ---------------------------------------------------------------------------
// Line pass along one axis for a tile of `tl` lines.
//   axis 0: lines are contiguous rows (line-major tile);
//   axis 1: line l = x + nx z, points strided by nx;
//   axis 2: line l = x + nx y, points strided by nx ny.
// For axes 1 and 2 consecutive lines are consecutive in memory, so the
// point-major tile is loaded with coalesced row segments.
// ---------------------------------------------------------------------------
template <typename R, int MODE, int TB, int MAXV>
__global__ __launch_bounds__(TB) void k_line(const PassDesc* __restrict__ descs, int step) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  if ((int)blockIdx.x >= d.ntiles) return;
  const int tid = threadIdx.x;
  const int axis = d.axis, nx = d.nx, ny = d.ny, nz = d.nz;
  const int n = d.p0.n;
  const long M = (long)nx * ny * nz;
  const int nlines = (int)(M / n);
  const int tl = d.tl;
  const int l0 = blockIdx.x * tl;
  const long S = axis == 0 ? 1 : (axis == 1 ? nx : (long)nx * ny);
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step;
  const cpx<R>* tw = reinterpret_cast<const cpx<R>*>(d.tw0);
  const R* g = reinterpret_cast<const R*>(d.g0);
  const int ntot = tl * n;
  int ls, ps;
  if (axis == 0) {
    ls = n;
    ps = 1;
    for (int idx = tid; idx < ntot; idx += TB) {
      const int l = idx / n, i = idx - l * n;
      const int gl = l0 + l;
      lds[idx] = (gl < nlines) ? in[(long)gl * n + i] : mk(R(0), R(0));
    }
  } else {
    ls = 1;
    ps = tl;
    for (int idx = tid; idx < ntot; idx += TB) {
      const int i = idx / tl, l = idx - i * tl;
      const int gl = l0 + l;
      long base;
      if (axis == 1) base = (long)(gl % nx) + (long)nx * ny * (gl / nx);
      else base = gl;
      lds[idx] = (gl < nlines) ? in[base + i * S] : mk(R(0), R(0));
    }
  }
  __syncthreads();
  if (MODE == MODE_CONV) {
    lds_fft<R, -1, MAXV>(d.p0, lds, tl, ls, ps, tw, g, tid, TB);
    lds_fft<R, +1, MAXV>(d.p0, lds, tl, ls, ps, tw, nullptr, tid, TB);
  } else if (MODE == MODE_FWD) {
    lds_fft<R, -1, MAXV>(d.p0, lds, tl, ls, ps, tw, nullptr, tid, TB);
  } else {
    lds_fft<R, +1, MAXV>(d.p0, lds, tl, ls, ps, tw, nullptr, tid, TB);
  }
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp);
    scale = R(d.scale);
  }
  for (int idx = tid; idx < ntot; idx += TB) {
    int l, i;
    if (axis == 0) {
      l = idx / n;
      i = idx - l * n;
    } else {
      i = idx / tl;
      l = idx - i * tl;
    }
    const int gl = l0 + l;
    if (gl >= nlines) continue;
    long base;
    if (axis == 0) base = (long)gl * n;
    else if (axis == 1) base = (long)(gl % nx) + (long)nx * ny * (gl / nx);
    else base = gl;
    const long gi = base + i * S;
    cpx<R> o = lds[idx];
    if (post) o = cmul(o, post[gi]);
    if (scale != R(1)) o = cscale(o, scale);
    out[gi] = o;
  }
}

// ---------------------------------------------------------------------------
// Field assembly, code/calc_poly_density.cpp:45-50 (AB pair):
//   wa = i w+ + (i wAB+ - wAB-)      [negative chi: i wAB- - wAB+]
//   wb = i w+ + (i wAB+ + wAB-)      [negative chi: -i wAB- - wAB+]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_assemble(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const double2 wp = s.wpl[i], p = s.wabp[i], m = s.wabm[i];
  const double2 iwp = mul_i(wp);
  double2 a, b;
  if (!s.neg_ab) {
    const double2 ip = mul_i(p);
    a = make_double2(iwp.x + (ip.x - m.x), iwp.y + (ip.y - m.y));
    b = make_double2(iwp.x + (ip.x + m.x), iwp.y + (ip.y + m.y));
  } else {
    const double2 im = mul_i(m);
    a = make_double2(iwp.x + (im.x - p.x), iwp.y + (im.y - p.y));
    b = make_double2(iwp.x + (-im.x - p.x), iwp.y + (-im.y - p.y));
  }
  // BC / AC pairs and the graft field, code/calc_poly_density.cpp:44-53:
  //   wa += i wAC+ - wAC-  [negative chi: i wAC- - wAC+]
  //   wb += i wBC+ - wBC-  [negative chi: i wBC- - wBC+]
  //   wg = i w+ + (i wAC+ + wAC-) + (i wBC+ + wBC-)
  //        [negative chi: -i wAC- - wAC+, -i wBC- - wBC+]
  double2 g = iwp;
  if (s.wacp) {
    const double2 cp = s.wacp[i], cm = s.wacm[i];
    if (!s.neg_ac) {
      const double2 icp = mul_i(cp);
      a = dadd(a, make_double2(icp.x - cm.x, icp.y - cm.y));
      g = dadd(g, make_double2(icp.x + cm.x, icp.y + cm.y));
    } else {
      const double2 icm = mul_i(cm);
      a = dadd(a, make_double2(icm.x - cp.x, icm.y - cp.y));
      g = dadd(g, make_double2(-icm.x - cp.x, -icm.y - cp.y));
    }
  }
  if (s.wbcp) {
    const double2 cp = s.wbcp[i], cm = s.wbcm[i];
    if (!s.neg_bc) {
      const double2 icp = mul_i(cp);
      b = dadd(b, make_double2(icp.x - cm.x, icp.y - cm.y));
      g = dadd(g, make_double2(icp.x + cm.x, icp.y + cm.y));
    } else {
      const double2 icm = mul_i(cm);
      b = dadd(b, make_double2(icm.x - cp.x, icm.y - cp.y));
      g = dadd(g, make_double2(-icm.x - cp.x, -icm.y - cp.y));
    }
  }
  s.wa[i] = a;
  s.wb[i] = b;
  if (s.wg) s.wg[i] = g;
}

// exp(-(smw - c)) into the propagator precision, and the chain ends
// q_0 = exp(-wA) (Na != 0) / q-dagger_0 = exp(-wB) (Na != N),
// code/diblock_discrete.cpp:19-29.
template <typename R>
__global__ __launch_bounds__(256) void k_exp_prop(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const double c = s.red[RED_SHIFT].y;
  const double2 a = s.wa[i], b = s.wb[i];
  const double2 ea = dexp(make_double2(-(a.x - c), -a.y));
  const double2 eb = dexp(make_double2(-(b.x - c), -b.y));
  cpx<R> ca = mk(R(ea.x), R(ea.y)), cb = mk(R(eb.x), R(eb.y));
  reinterpret_cast<cpx<R>*>(s.expA)[i] = ca;
  reinterpret_cast<cpx<R>*>(s.expB)[i] = cb;
  reinterpret_cast<cpx<R>*>(s.q)[i] = (s.Na != 0) ? ca : cb;
  if (s.qd != s.q) reinterpret_cast<cpx<R>*>(s.qd)[i] = (s.Na != s.N) ? cb : ca;
  if (s.grafted) {
    // free-end graft propagator q_g[0] = exp(-smwg),
    // code/grafted_nanoparticle.cpp:24-26
    const double2 g = s.wg[i];
    const double2 eg = dexp(make_double2(-(g.x - c), -g.y));
    const cpx<R> cg = mk(R(eg.x), R(eg.y));
    reinterpret_cast<cpx<R>*>(s.expG)[i] = cg;
    reinterpret_cast<cpx<R>*>(s.qg)[i] = cg;
  }
}

// complementary graft propagator start, code/grafted_nanoparticle.cpp:55-57:
// q_g-dagger[0] = expl_grafts exp(-smwg) / q_g[Ng-1]
template <typename R>
__global__ __launch_bounds__(256) void k_graft_init(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (!s.grafted || i >= s.M) return;
  const cpx<R> e = reinterpret_cast<const cpx<R>*>(s.expG)[i];
  const cpx<R> ql = reinterpret_cast<const cpx<R>*>(s.qg)[(long)(s.Ng - 1) * s.M + i];
  const double2 num = dmul(s.expl[i], make_double2((double)e.x, (double)e.y));
  const double2 v = ddiv(num, make_double2((double)ql.x, (double)ql.y));
  reinterpret_cast<cpx<R>*>(s.qgd)[i] = mk(R(v.x), R(v.y));
}

// graft density, code/grafted_nanoparticle.cpp:73-80: rhog_exp =
// ng_per_np n_plates exp(smwg) sum_j q_g[j] q_g-dagger[Ng-1-j] (FP64
// accumulation; with the gauge shift the product carries exp(c))
template <typename R>
__global__ __launch_bounds__(256) void k_graft_density(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (!s.grafted || i >= s.M) return;
  const cpx<R>* q = reinterpret_cast<const cpx<R>*>(s.qg);
  const cpx<R>* qd = reinterpret_cast<const cpx<R>*>(s.qgd);
  const long M = s.M;
  double2 acc = make_double2(0.0, 0.0);
  for (int j = 0; j < s.Ng; j++) {
    const cpx<R> a = q[(long)j * M + i];
    const cpx<R> b = qd[(long)(s.Ng - 1 - j) * M + i];
    acc = dadd(acc, make_double2((double)a.x * b.x - (double)a.y * b.y, (double)a.x * b.y + (double)a.y * b.x));
  }
  const double c = s.red[RED_SHIFT].y;
  const double2 g = s.wg[i];
  const double2 e = dexp(make_double2(g.x - c, g.y));
  s.rhog[i] = dscale(dmul(acc, e), s.ng_per_np * (double)s.n_exp);
}

// Graft sites (code/initialize.cpp:160-196, field_gradient at
// code/array_utils.cpp:565-579): out = i k_d in (spectrum), k_d the aliased
// wavenumber (0 at the Nyquist mode of an even axis)
__global__ __launch_bounds__(256) void k_ik_mult(const SysDev* __restrict__ sys, const double2* __restrict__ in,
                                                 double2* __restrict__ out, int d) {
  const SysDev& s = sys[0];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const int n[3] = {(int)(i % s.nx), (int)((i / s.nx) % s.ny), (int)(i / ((long)s.nx * s.ny))};
  const int Ns[3] = {s.nx, s.ny, s.nz};
  double k;
  if (Ns[d] % 2 == 0 && n[d] == Ns[d] / 2) k = 0.0;
  else if ((double)n[d] < (double)Ns[d] / 2.0) k = 2.0 * 3.14159265358979323846 * n[d] / s.Lb[d];
  else k = 2.0 * 3.14159265358979323846 * (n[d] - Ns[d]) / s.Lb[d];
  out[i] = mul_i(dscale(in[i], k));
}
// mode 0: acc += a * a;  mode 1: acc2 += sqrt(acc) (principal branch)
__global__ __launch_bounds__(256) void k_sites_accum(long M, const double2* __restrict__ a, double2* __restrict__ acc,
                                                     int mode) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  const double2 v = a[i];
  if (mode == 0) {
    acc[i] = dadd(acc[i], dmul(v, v));
  } else {
    const double r = sqrt(v.x * v.x + v.y * v.y);
    const double re = sqrt(0.5 * (r + v.x));
    const double im = copysign(sqrt(0.5 * fmax(r - v.x, 0.0)), v.y);
    acc[i] = dadd(acc[i], make_double2(re, im));
  }
}

// ---------------------------------------------------------------------------
// Deterministic two-stage FP64 reductions.  Stage 1: block b of system s
// sums points b*256+t, stride nblk*256, then a fixed LDS tree.  Stage 2: one
// block sums the nblk partials in a fixed order into s.red[slot].
// ---------------------------------------------------------------------------
template <int WHAT, typename R>
__device__ __forceinline__ double2 red_term(const SysDev& s, long i) {
  if constexpr (WHAT == RED_Q) {
    const cpx<R> v = ld<R>(s.qlast, i);
    return make_double2((double)v.x, (double)v.y);
  } else if constexpr (WHAT == RED_H) {
    // wpl_part + wab_part integrands, code/calc_h.cpp:70-73, :97-98
    const double2 w = s.wpl[i];
    double2 t = make_double2(0.0, 0.0);
    if (s.kappa > 0.0) {
      const double2 w2 = dmul(w, w);
      t = dscale(w2, s.rho0 / s.kappa / 2.0);
    }
    // rho0 (1 - rho_np), rho_np stored complex
    const double2 c = make_double2(s.rho0 * (1.0 - s.rho_np[i].x), -s.rho0 * s.rho_np[i].y);
    const double2 icw = mul_i(dmul(c, w));
    t = make_double2(t.x - icw.x, t.y - icw.y);
    // wbc_part, wac_part, wab_part (code/calc_h.cpp:89-132), in the order
    // of the sum at :13
    if (s.wbcp && s.chi_bc != 0.0) {
      const double2 p = s.wbcp[i], m = s.wbcm[i];
      t = dadd(t, dscale(dadd(dmul(p, p), dmul(m, m)), s.rho0 / s.chi_bc * (s.neg_bc ? -1.0 : 1.0)));
    }
    if (s.wacp && s.chi_ac != 0.0) {
      const double2 p = s.wacp[i], m = s.wacm[i];
      t = dadd(t, dscale(dadd(dmul(p, p), dmul(m, m)), s.rho0 / s.chi_ac * (s.neg_ac ? -1.0 : 1.0)));
    }
    if (s.chi_ab != 0.0) {
      const double2 p = s.wabp[i], m = s.wabm[i];
      const double2 pp = dmul(p, p), mm = dmul(m, m);
      const double sg = s.neg_ab ? -1.0 : 1.0;
      t = dadd(t, dscale(dadd(pp, mm), s.rho0 / s.chi_ab * sg));
    }
    return t;
  } else if constexpr (WHAT == RED_SHIFT) {
    return make_double2(s.fA * s.wa[i].x + (1.0 - s.fA) * s.wb[i].x, 0.0);
  } else if constexpr (WHAT == RED_VEXP) {
    return s.rho_np[i];
  } else if constexpr (WHAT == RED_QGA) {
    // code/grafted_nanoparticle.cpp:83-84: expl_grafts log q_g[Ng-1] where
    // Re expl_grafts > 0
    if (!s.grafted) return make_double2(0.0, 0.0);
    const double2 e = s.expl[i];
    if (e.x <= 0.0) return make_double2(0.0, 0.0);
    const cpx<R> v = ld<R>(s.qg, (long)(s.Ng - 1) * s.M + i);
    const double vx = v.x, vy = v.y;
    const double2 lg = make_double2(0.5 * log(vx * vx + vy * vy), atan2(vy, vx));
    return dmul(e, lg);
  } else {
    return dadd(s.rhoA[i], s.rhoB[i]);
  }
}

template <int WHAT, typename R>
__global__ __launch_bounds__(256) void k_reduce1(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  __shared__ double2 sh[256];
  double2 acc = make_double2(0.0, 0.0);
  if ((int)blockIdx.x < s.nblk) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < s.M; i += (long)s.nblk * 256)
      acc = dadd(acc, red_term<WHAT, R>(s, i));
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = dadd(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && (int)blockIdx.x < s.nblk) s.partials[WHAT * s.nblk + blockIdx.x] = sh[0];
}

template <int WHAT>
__global__ __launch_bounds__(256) void k_reduce2(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  __shared__ double2 sh[256];
  double2 acc = make_double2(0.0, 0.0);
  for (int b = threadIdx.x; b < s.nblk; b += 256) acc = dadd(acc, s.partials[WHAT * s.nblk + b]);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = dadd(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double2 r = sh[0];
    if (WHAT == RED_SHIFT) r = make_double2(r.x, r.x / (double)s.M);  // .y = mean
    s.red[WHAT] = r;
  }
}

// ---------------------------------------------------------------------------
// Density, code/calc_poly_density.cpp:167-191:
//   rhoA = nD/(Q V) exp(smwA) sum_{j<Na} q_j qdag_{N-1-j}, rhoB over j >= Na,
// accumulated in FP64 whatever the propagator precision.  With the gauge
// shift c both Q and the q's carry exp(N c); exp(smw - c) restores rho.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_density(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const double2 Qs = dscale(s.red[RED_Q], s.dV / s.V);
  const double2 factor = ddiv(make_double2(s.nD, 0.0), dscale(Qs, s.V));
  const double c = s.red[RED_SHIFT].y;
  double2 accA = make_double2(0.0, 0.0), accB = make_double2(0.0, 0.0);
  const cpx<R>* q = reinterpret_cast<const cpx<R>*>(s.q);
  const cpx<R>* qd = reinterpret_cast<const cpx<R>*>(s.qd);
  const long M = s.M;
  for (int j = 0; j < s.N; j++) {
    const cpx<R> a = q[(long)j * M + i];
    const cpx<R> b = qd[(long)(s.N - 1 - j) * M + i];
    const double2 p = make_double2((double)a.x * b.x - (double)a.y * b.y,
                                   (double)a.x * b.y + (double)a.y * b.x);
    if (s.Na > 0 && j < s.Na) accA = dadd(accA, p);
    else accB = dadd(accB, p);
  }
  const double2 wa = s.wa[i], wb = s.wb[i];
  const double2 ea = dexp(make_double2(wa.x - c, wa.y));
  const double2 eb = dexp(make_double2(wb.x - c, wb.y));
  s.rhoA[i] = dmul(accA, dmul(ea, factor));
  s.rhoB[i] = dmul(accB, dmul(eb, factor));
}

// Density of one segment of the checkpointed propagators: q_j for
// j = lo .. lo + n - 1 in slices 1..n of the q window, q-dagger_{N-1-j} in
// slice N - 1 - j - qd0 + 1 of the q-dagger window (qd0 = first q-dagger
// index of the window).  The first segment assigns, later ones add; the last
// applies nD / (Q V) exp(smw - c) like k_density.
template <typename R>
__global__ __launch_bounds__(256) void k_density_seg(const SysDev* __restrict__ sys, int lo, int n, int qd0,
                                                     int first, int last) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const cpx<R>* q = reinterpret_cast<const cpx<R>*>(s.ck_sq);
  const cpx<R>* qd = reinterpret_cast<const cpx<R>*>(s.ck_tq);
  const long M = s.M;
  double2 accA = make_double2(0.0, 0.0), accB = make_double2(0.0, 0.0);
  for (int u = 1; u <= n; u++) {
    const int j = lo + u - 1;
    const cpx<R> a = q[(long)u * M + i];
    const cpx<R> b = qd[(long)(s.N - 1 - j - qd0 + 1) * M + i];
    const double2 p = make_double2((double)a.x * b.x - (double)a.y * b.y,
                                   (double)a.x * b.y + (double)a.y * b.x);
    if (s.Na > 0 && j < s.Na) accA = dadd(accA, p);
    else accB = dadd(accB, p);
  }
  if (!first) {
    accA = dadd(accA, s.rhoA[i]);
    accB = dadd(accB, s.rhoB[i]);
  }
  if (last) {
    const double2 Qs = dscale(s.red[RED_Q], s.dV / s.V);
    const double2 factor = ddiv(make_double2(s.nD, 0.0), dscale(Qs, s.V));
    const double c = s.red[RED_SHIFT].y;
    const double2 wa = s.wa[i], wb = s.wb[i];
    accA = dmul(accA, dmul(dexp(make_double2(wa.x - c, wa.y)), factor));
    accB = dmul(accB, dmul(dexp(make_double2(wb.x - c, wb.y)), factor));
  }
  s.rhoA[i] = accA;
  s.rhoB[i] = accB;
}

// Copy of one slice (M propagator values) between the checkpoint arrays of
// every system: array 0 = q window, 1 = q-dagger window, 2 = checkpoints;
// an optional second destination (dst2 < 0: none).  n16 = 16-byte units
// per slice.
__device__ __forceinline__ float4* ck_slice(const SysDev& s, int arr, int slot, int esz) {
  char* base = (char*)(arr == 0 ? s.ck_sq : arr == 1 ? s.ck_tq : s.ck_cq);
  return reinterpret_cast<float4*>(base + (size_t)slot * s.M * esz);
}
__global__ __launch_bounds__(256) void k_ck_copy(const SysDev* __restrict__ sys, int esz, int src, int src_slot,
                                                 int dst, int dst_slot, int dst2, int dst2_slot) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)s.M * esz / 16) return;
  const float4 v = ck_slice(s, src, src_slot, esz)[i];
  ck_slice(s, dst, dst_slot, esz)[i] = v;
  if (dst2 >= 0) ck_slice(s, dst2, dst2_slot, esz)[i] = v;
}

// ---------------------------------------------------------------------------
// 1S semi-implicit update in k-space, code/1s_update.cpp:57-160 (w+ and the
// AB pair).  The k = 0 delta of the w+ force is at linear index 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_update_1s(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const double rho0 = s.rho0, li = s.lam_imag, lr = s.lam_real;
  const double h = s.hhat_k[i];
  const double2 rA = s.rhoA[i], rB = s.rhoB[i];
  const double2 rsum = dadd(rA, rB);
  const double2 enr = s.exp_nrH[i];
  const double evar = (i == 0) ? 1.0 : 0.0;
  const double2 rg = s.grafted ? s.rhog[i] : make_double2(0.0, 0.0);  // k-space
  {
    // F = i rho0 (exp_nrH - evar) + i h (rhoA + rhog + rhoB)
    const double2 all = dadd(dadd(rA, rg), rB);
    const double2 in = make_double2(rho0 * (enr.x - evar) + h * all.x, rho0 * enr.y + h * all.y);
    const double2 F = mul_i(in);
    double A = (s.kappa <= 0.0 ? 0.0 : rho0 / s.kappa) + s.Xk[i];
    if (s.grafted) A += s.XGk[i];
    const double B = s.kappa > 0.0 ? A - rho0 / s.kappa : 0.0;
    const double2 w = s.wpl[i];
    double2 num = make_double2(w.x - li * (F.x - B * w.x), w.y - li * (F.y - B * w.y));
    if (s.do_cl) num = dadd(num, s.etap[i]);
    s.wpl[i] = dscale(num, 1.0 / (1.0 + li * A));
  }
  if (s.chi_ab != 0.0) {
    const double2 wp = s.wabp[i], wm = s.wabm[i];
    const double c2 = 2.0 * rho0 / s.chi_ab;
    if (!s.neg_ab) {
      double2 F = mul_i(dscale(rsum, h));
      if (s.np_chem == 1) F = make_double2(F.x - rho0 * enr.x, F.y - rho0 * enr.y);
      const double A = c2 + s.Xk[i];
      const double B = A - c2;
      double2 num = make_double2(wp.x - li * (F.x - B * wp.x), wp.y - li * (F.y - B * wp.y));
      if (s.do_cl) num = dadd(num, s.etap[s.M + i]);
      s.wabp[i] = dscale(num, 1.0 / (1.0 + li * A));
      double2 Fm = dscale(make_double2(rB.x - rA.x, rB.y - rA.y), h);
      if (s.np_chem == 1) Fm = dadd(Fm, mul_i(dscale(enr, rho0)));
      double2 numm = make_double2(wm.x - lr * Fm.x, wm.y - lr * Fm.y);
      if (s.do_cl) numm = dadd(numm, s.etam[i]);
      s.wabm[i] = dscale(numm, 1.0 / (1.0 + lr * c2));
    } else {
      const double2 F = dscale(rsum, -h);
      s.wabp[i] = dscale(make_double2(wp.x - lr * F.x, wp.y - lr * F.y), 1.0 / (1.0 + lr * c2));
      double2 Fm = mul_i(dscale(make_double2(rA.x - rB.x, rA.y - rB.y), h));
      if (s.np_chem == 1) Fm = make_double2(Fm.x - rho0 * enr.x, Fm.y - rho0 * enr.y);
      const double A = c2 + s.X2k[i];
      const double B = A - c2;
      double2 num = make_double2(wm.x - li * (Fm.x - B * wm.x), wm.y - li * (Fm.y - B * wm.y));
      if (s.do_cl) num = dadd(num, s.etap[s.M + i]);
      s.wabm[i] = dscale(num, 1.0 / (1.0 + li * A));
    }
  } else {
    s.wabp[i] = make_double2(0.0, 0.0);
    s.wabm[i] = make_double2(0.0, 0.0);
  }
  // BC pair (rho = rhoB, g = gbb; code/1s_update.cpp:162-247) and AC pair
  // (rho = rhoA, g = gaa; :249-329), both signs of chi
  for (int pr = 0; pr < 2; pr++) {
    double2* wpp = pr == 0 ? s.wbcp : s.wacp;
    double2* wmp = pr == 0 ? s.wbcm : s.wacm;
    if (!wpp) continue;
    const double chi = pr == 0 ? s.chi_bc : s.chi_ac;
    if (chi == 0.0) {
      wpp[i] = make_double2(0.0, 0.0);
      wmp[i] = make_double2(0.0, 0.0);
      continue;
    }
    const bool neg = pr == 0 ? s.neg_bc : s.neg_ac;
    const double2 rho = pr == 0 ? rB : rA;
    const double G = pr == 0 ? s.XBk[i] : s.XAk[i];
    const double GC = s.grafted ? s.XGk[i] : 0.0;
    const double c2 = 2.0 * rho0 / chi;
    const double2 wp = wpp[i], wm = wmp[i];
    if (!neg) {
      // +: F = i h (rhog + rho), A = 2 rho0/chi + G + GC, semi-implicit
      const double2 F = mul_i(dscale(dadd(rg, rho), h));
      const double A = c2 + G + GC;
      const double B = A - c2;
      wpp[i] = dscale(make_double2(wp.x - li * (F.x - B * wp.x), wp.y - li * (F.y - B * wp.y)),
                      1.0 / (1.0 + li * A));
      // -: F = h (rhog - rho) (+ i rho0 exp_nrH for A-like plates), explicit
      double2 Fm = dscale(make_double2(rg.x - rho.x, rg.y - rho.y), h);
      if (s.np_chem == 1) Fm = dadd(Fm, mul_i(dscale(enr, rho0)));
      wmp[i] = dscale(make_double2(wm.x - lr * Fm.x, wm.y - lr * Fm.y), 1.0 / (1.0 + lr * c2));
    } else {
      // +: F = -h (rho + rhog), explicit
      const double2 F = dscale(dadd(rho, rg), -h);
      wpp[i] = dscale(make_double2(wp.x - lr * F.x, wp.y - lr * F.y), 1.0 / (1.0 + lr * c2));
      // -: F = -i h (rhog - rho) (- rho0 exp_nrH), A = 2 rho0/chi + G - GC
      double2 Fm = mul_i(dscale(make_double2(rg.x - rho.x, rg.y - rho.y), -h));
      if (s.np_chem == 1) Fm = make_double2(Fm.x - rho0 * enr.x, Fm.y - rho0 * enr.y);
      const double A = c2 + G - GC;
      const double B = A - c2;
      wmp[i] = dscale(make_double2(wm.x - li * (Fm.x - B * wm.x), wm.y - li * (Fm.y - B * wm.y)),
                      1.0 / (1.0 + li * A));
    }
  }
}

// Explicit Euler / complex-Langevin step in k-space,
// code/Euler_update.cpp:52-99 (w+ and the AB pair).
__global__ __launch_bounds__(256) void k_update_euler(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const double rho0 = s.rho0, li = s.lam_imag, lr = s.lam_real;
  const double h = s.hhat_k[i];
  const double2 rA = s.rhoA[i], rB = s.rhoB[i];
  const double2 enr = s.exp_nrH[i];
  const double evar = (i == 0) ? 1.0 : 0.0;
  {
    const double2 w = s.wpl[i];
    double2 F = s.kappa <= 0.0 ? make_double2(0.0, 0.0) : dscale(w, rho0 / s.kappa);
    const double2 in = make_double2(rho0 * (enr.x - evar) + h * (rA.x + rB.x),
                                    rho0 * enr.y + h * (rA.y + rB.y));
    F = dadd(F, mul_i(in));
    double2 num = make_double2(w.x - li * F.x, w.y - li * F.y);
    if (s.do_cl) num = dadd(num, s.etap[i]);
    s.wpl[i] = num;
  }
  if (s.chi_ab != 0.0) {
    const double2 wp = s.wabp[i], wm = s.wabm[i];
    const double c2 = 2.0 * rho0 / s.chi_ab;
    // A segments and B segments enter with sp (AB+) and sa, sb (AB-)
    double2 Fp, Fm;
    if (!s.neg_ab) {
      Fp = dadd(dscale(wp, c2), mul_i(dscale(dadd(rA, rB), h)));
      Fm = dadd(dscale(wm, c2), dscale(make_double2(rB.x - rA.x, rB.y - rA.y), h));
    } else {
      Fp = dadd(dscale(wp, c2), dscale(dadd(rA, rB), -h));
      Fm = dadd(dscale(wm, c2), mul_i(dscale(make_double2(rA.x - rB.x, rA.y - rB.y), h)));
    }
    double2 np_ = make_double2(wp.x - li * Fp.x, wp.y - li * Fp.y);
    double2 nm = make_double2(wm.x - lr * Fm.x, wm.y - lr * Fm.y);
    if (s.do_cl) {
      np_ = dadd(np_, s.etap[s.M + i]);
      nm = dadd(nm, s.etam[i]);
    }
    s.wabp[i] = np_;
    s.wabm[i] = nm;
  } else {
    s.wabp[i] = make_double2(0.0, 0.0);
    s.wabm[i] = make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Complex-Langevin noise, generate_1s_noise (code/1s_update.cpp:356-368):
// real Gaussian noise of scale sqrt(2 lambda) / prod_d sqrt(dx_d) in real
// space (forward-transformed afterwards).  Counter-based Philox4x32-10
// (Salmon et al., SC'11) keyed by the system's seed; counter = (point,
// field, iteration): every point draws independently, a replayed graph gets
// a new iteration number from the device-side counter, and a system's noise
// does not depend on its place in a batch.  Box-Muller on two 53-bit
// uniforms.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox_round(uint32_t c[4], uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k[2]) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    philox_round(c, k);
    if (r < 9) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
  }
}

__global__ __launch_bounds__(256) void k_noise(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (!s.do_cl || i >= s.M) return;
  const unsigned long long it = *s.rng_ctr;
#pragma unroll
  for (int f = 0; f < 3; f++) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)f, (uint32_t)it, (uint32_t)(it >> 32)};
    uint32_t k[2] = {(uint32_t)s.seed, (uint32_t)(s.seed >> 32)};
    philox4x32_10(c, k);
    const double u1 = ((double)((((uint64_t)c[0]) << 21) ^ (uint64_t)c[1]) + 0.5) * 0x1p-53;
    const double u2 = ((double)((((uint64_t)c[2]) << 21) ^ (uint64_t)c[3]) + 0.5) * 0x1p-53;
    const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
    const double2 v = make_double2(s.noise_scale[f] * z, 0.0);
    if (f == 0) s.etap[i] = v;
    else if (f == 1) s.etap[s.M + i] = v;
    else s.etam[i] = v;
  }
}

__global__ void k_noise_advance(const SysDev* __restrict__ sys, int nsys) {
  const int j = threadIdx.x + blockIdx.x * blockDim.x;
  if (j < nsys && sys[j].do_cl) *sys[j].rng_ctr += 1;
}

// Complex-Langevin averages after the update of iteration it = counter - 1
// (the counter was advanced when that iteration drew its noise):
// accumulate when it >= sample_wait and it % sample_freq == 0
// (code/simulate.cpp:91-92, code/array_utils.cpp:13-35).
__global__ __launch_bounds__(256) void k_cl_accumulate(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  if (!s.do_cl) return;
  const long long it = (long long)*s.rng_ctr - 1;
  if (it < s.sample_wait || it % s.sample_freq != 0) return;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < s.M) {
    const double2 a = s.rhoA[i], b = s.rhoB[i];
    double2 sa = s.avgA[i], sb = s.avgB[i];
    sa.x += a.x;
    sa.y += a.y;
    sb.x += b.x;
    sb.y += b.y;
    s.avgA[i] = sa;
    s.avgB[i] = sb;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *s.n_samples += 1.0;
}

// ---------------------------------------------------------------------------
// Plate density and its derivatives.  Rod (np_type 2, explicit_nanorod,
// code/initialize.cpp:578-612): 1/4 erfc(a) erfc(b), a = (|u.dr| - len/2)/xi,
// b = (|u x dr| - R)/xi, dr the minimum image of x - c (x = n dx).  Square
// plate (np_type 3, 3D): 1/8 erfc(a) erfc(b1) erfc(b2), b_i = (|e_i.dr| - R)/xi.
// Derivatives are those of the sampled formula: d/dc (dr -> -1) and the
// rotation of the plate's axes about `axis` through its centre
// (d v / d theta = axis x v).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
// d erfc(z) / dz
__device__ __forceinline__ double derfc(double z) { return -1.1283791670955125739 * exp(-z * z); }

// `scale` inflates the length and radius (1.05 for the graft sites)
template <bool GRAD>
__device__ __forceinline__ double plate_eval(const SysDev& s, int p, long i, const double* axis, double* dc,
                                             double* dth, double scale = 1.0) {
  const int nn[3] = {(int)(i % s.nx), (int)((i / s.nx) % s.ny), (int)(i / ((long)s.nx * s.ny))};
  const PlateDev& P = s.plate[p];
  double dr[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < s.dim; d++) {
    dr[d] = (double)nn[d] * s.dxb[d] - P.c[d];
    if (dr[d] >= 0.5 * s.Lb[d]) dr[d] -= s.Lb[d];
    else if (dr[d] < -0.5 * s.Lb[d]) dr[d] += s.Lb[d];
  }
  const double xi = s.xi_nr, len = scale * s.len_nr, rad = scale * s.R_nr;
  const double ud = dot3(P.u, dr);
  const double a = (fabs(ud) - 0.5 * len) / xi;
  const double ea = erfc(a);
  const double sud = ud > 0.0 ? 1.0 : (ud < 0.0 ? -1.0 : 0.0);
  double du[3] = {0.0, 0.0, 0.0};
  if (GRAD) cross3(axis, P.u, du);
  if (s.np_type == SCFT_NP_SQUARE && s.dim == 3) {
    const double a1 = dot3(P.e1, dr), a2 = dot3(P.e2, dr);
    const double b1 = (fabs(a1) - rad) / xi, b2 = (fabs(a2) - rad) / xi;
    const double e1v = erfc(b1), e2v = erfc(b2);
    if (GRAD) {
      const double s1 = a1 > 0.0 ? 1.0 : (a1 < 0.0 ? -1.0 : 0.0);
      const double s2 = a2 > 0.0 ? 1.0 : (a2 < 0.0 ? -1.0 : 0.0);
      const double ga = 0.125 * derfc(a) * e1v * e2v / xi;
      const double g1 = 0.125 * ea * derfc(b1) * e2v / xi;
      const double g2 = 0.125 * ea * e1v * derfc(b2) / xi;
      double de1[3], de2[3];
      cross3(axis, P.e1, de1);
      cross3(axis, P.e2, de2);
      for (int k = 0; k < 3; k++) dc[k] = -(ga * sud * P.u[k] + g1 * s1 * P.e1[k] + g2 * s2 * P.e2[k]);
      *dth = ga * sud * dot3(du, dr) + g1 * s1 * dot3(de1, dr) + g2 * s2 * dot3(de2, dr);
    }
    return 0.125 * ea * e1v * e2v;
  }
  double sc, dsc_dr[3] = {0.0, 0.0, 0.0}, dsc_dth = 0.0;
  if (s.dim == 2) {
    const double cz = P.u[0] * dr[1] - P.u[1] * dr[0];
    sc = fabs(cz);
    if (GRAD) {
      const double sg = cz > 0.0 ? 1.0 : (cz < 0.0 ? -1.0 : 0.0);
      dsc_dr[0] = -sg * P.u[1];
      dsc_dr[1] = sg * P.u[0];
      dsc_dth = sg * (du[0] * dr[1] - du[1] * dr[0]);
    }
  } else {
    double cr[3];
    cross3(P.u, dr, cr);
    sc = sqrt(dot3(cr, cr));
    if (GRAD && sc > 0.0) {
      for (int k = 0; k < 3; k++) dsc_dr[k] = (dr[k] - ud * P.u[k]) / sc;
      dsc_dth = -ud * dot3(du, dr) / sc;
    }
  }
  const double b = (sc - rad) / xi;
  const double eb = erfc(b);
  if (GRAD) {
    const double ga = 0.25 * derfc(a) * eb / xi;
    const double gb = 0.25 * ea * derfc(b) / xi;
    for (int k = 0; k < 3; k++) dc[k] = -(ga * sud * P.u[k] + gb * dsc_dr[k]);
    *dth = ga * sud * dot3(du, dr) + gb * dsc_dth;
  }
  return 0.25 * ea * eb;
}

__global__ __launch_bounds__(256) void k_cavity(const SysDev* __restrict__ sys) {
  const SysDev& s = sys[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  double rho = 0.0;
  for (int p = 0; p < s.n_plates; p++) rho += plate_eval<false>(s, p, i, nullptr, nullptr, nullptr);
  s.rho_np[i] = make_double2(rho, 0.0);
}

// sum over points of d rho_p / d g times W = i rho0 w+ + (rho0 / N) log Q,
// times dV: the explicit derivative of H (code/calc_h.cpp:71-72 through
// rho_np, and -nD log Q through nD = rho0 (Vf - V_exp) / N)
__global__ __launch_bounds__(256) void k_plate_grad1(const SysDev* __restrict__ sysp, double ax, double ay,
                                                     double az, double lq_re, double lq_im) {
  const SysDev& s = *sysp;
  const double axis[3] = {ax, ay, az};
  __shared__ double2 sh[8][256];
  double2 acc[8];
  for (int k = 0; k < 8; k++) acc[k] = make_double2(0.0, 0.0);
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < s.M; i += (long)s.nblk * 256) {
    const double2 w = s.wpl[i];
    // W = i rho0 w + (rho0 / N) log Q
    const double2 W = make_double2(-s.rho0 * w.y + s.rho0 / s.N * lq_re, s.rho0 * w.x + s.rho0 / s.N * lq_im);
    for (int p = 0; p < s.n_plates; p++) {
      double dc[3], dth;
      plate_eval<true>(s, p, i, axis, dc, &dth);
      for (int k = 0; k < 3; k++) acc[4 * p + k] = dadd(acc[4 * p + k], dscale(W, dc[k]));
      acc[4 * p + 3] = dadd(acc[4 * p + 3], dscale(W, dth));
    }
  }
  for (int k = 0; k < 8; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 8; k++) sh[k][threadIdx.x] = dadd(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 8) s.grad_partials[(long)threadIdx.x * s.nblk + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_plate_grad2(const SysDev* __restrict__ sysp, double2* out) {
  const SysDev& s = *sysp;
  __shared__ double2 sh[256];
  for (int k = 0; k < 8; k++) {
    double2 acc = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < s.nblk; b += 256) acc = dadd(acc, s.grad_partials[(long)k * s.nblk + b]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] = dadd(sh[threadIdx.x], sh[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = dscale(sh[0], s.dV);
    __syncthreads();
  }
}

// Graft-site derivatives (the plate dependence of -n ng_per_np Re Qga):
// the inflated shape of plate p (lam < 0) or its derivative with respect to
// centre coordinate lam (0..2) or to a rotation about `axis` (lam 3)
__global__ __launch_bounds__(256) void k_site_shape(const SysDev* __restrict__ sysp, int p, int lam, double ax,
                                                    double ay, double az, double scale, double2* __restrict__ out) {
  const SysDev& s = *sysp;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= s.M) return;
  const double axis[3] = {ax, ay, az};
  double dc[3], dth, v;
  if (lam < 0) {
    v = plate_eval<false>(s, p, i, nullptr, nullptr, nullptr, scale);
  } else {
    plate_eval<true>(s, p, i, axis, dc, &dth, scale);
    v = lam < 3 ? dc[lam] : dth;
  }
  out[i] = make_double2(v, 0.0);
}

// acc += a b
__global__ __launch_bounds__(256) void k_cmul_acc(long M, const double2* __restrict__ a,
                                                  const double2* __restrict__ b, double2* __restrict__ acc) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < M) acc[i] = dadd(acc[i], dmul(a[i], b[i]));
}

// block partial sums of (S / g) (log q_g[Ng-1] - Ng c - Qga) over the points
// where Re expl_grafts > 0 (the points Qga sums over); S = sum_d G_d G'_d, so
// S / g is the derivative of g = |grad shape|.  Where g vanishes (|g|^2 <=
// g2_min: the plate centre and symmetry lines, at round-off level) g has a
// kink whose one-sided derivatives are +-|G'|; those points add 0, the
// symmetric derivative
template <typename R>
__global__ __launch_bounds__(256) void k_site_grad_partial(const SysDev* __restrict__ sysp,
                                                           const double2* __restrict__ S,
                                                           const double2* __restrict__ g, double q_re, double q_im,
                                                           double g2_min, double2* __restrict__ partials) {
  const SysDev& s = *sysp;
  __shared__ double2 sh[256];
  double2 acc = make_double2(0.0, 0.0);
  const double c = s.red[RED_SHIFT].y;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < s.M; i += (long)gridDim.x * 256) {
    const double2 gv = g[i];
    const double gg = gv.x * gv.x + gv.y * gv.y;
    if (s.expl[i].x <= 0.0 || gg <= g2_min) continue;
    const cpx<R> v = ld<R>(s.qg, (long)(s.Ng - 1) * s.M + i);
    const double vx = v.x, vy = v.y;
    const double2 L = make_double2(0.5 * log(vx * vx + vy * vy) - (double)s.Ng * c - q_re, atan2(vy, vx) - q_im);
    acc = dadd(acc, dmul(ddiv(S[i], gv), L));
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = dadd(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_scale_real(double* a, long n, double f) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] *= f;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
template <typename R, int MODE, int TB, bool PLANE>
static void set_lds_attr() {
  static bool done[64] = {};
  if (PLANE)
    raise_lds_limit((const void*)k_plane<R, MODE, TB, 32>, 160 * 1024, done);
  else
    raise_lds_limit((const void*)k_line<R, MODE, TB, 32>, 160 * 1024, done);
}

template <typename R, int MODE>
static hipError_t launch_pass_t(bool plane, const PassDesc* d, int njobs, int gridx, int tb,
                                size_t lds, int step, hipStream_t st) {
  dim3 grid(gridx, njobs);
#define SCFT_LAUNCH(TB)                                                                  \
  do {                                                                                   \
    if (plane) {                                                                         \
      if (lds > 64 * 1024) set_lds_attr<R, MODE, TB, true>();                             \
      hipLaunchKernelGGL((k_plane<R, MODE, TB, 32>), grid, dim3(TB), lds, st, d, step);  \
    } else {                                                                             \
      if (lds > 64 * 1024) set_lds_attr<R, MODE, TB, false>();                            \
      hipLaunchKernelGGL((k_line<R, MODE, TB, 32>), grid, dim3(TB), lds, st, d, step);   \
    }                                                                                    \
  } while (0)
  if (tb <= 256) SCFT_LAUNCH(256);
  else if (tb <= 512) SCFT_LAUNCH(512);
  else SCFT_LAUNCH(1024);
#undef SCFT_LAUNCH
  return hipGetLastError();
}

hipError_t launch_pass(int prec, int mode, bool plane, const PassDesc* d, int njobs, int gridx,
                       int tb, size_t lds, int step, hipStream_t st) {
  if (prec == 0) {
    if (mode == MODE_CONV) return launch_pass_t<double, MODE_CONV>(plane, d, njobs, gridx, tb, lds, step, st);
    if (mode == MODE_FWD) return launch_pass_t<double, MODE_FWD>(plane, d, njobs, gridx, tb, lds, step, st);
    return launch_pass_t<double, MODE_INV>(plane, d, njobs, gridx, tb, lds, step, st);
  }
  if (mode == MODE_CONV) return launch_pass_t<float, MODE_CONV>(plane, d, njobs, gridx, tb, lds, step, st);
  if (mode == MODE_FWD) return launch_pass_t<float, MODE_FWD>(plane, d, njobs, gridx, tb, lds, step, st);
  return launch_pass_t<float, MODE_INV>(plane, d, njobs, gridx, tb, lds, step, st);
}

static inline dim3 pgrid(int maxM, int nsys) { return dim3((maxM + 255) / 256, nsys); }

// Plain 16-byte copy: the calibration kernel of known traffic for the
// hardware byte counters (read n16 * 16 bytes, write n16 * 16 bytes).
__global__ __launch_bounds__(256) void k_copy16(const float4* __restrict__ a, float4* __restrict__ b, long n16) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n16; i += (long)gridDim.x * 256) b[i] = a[i];
}

hipError_t launch_copy16(const void* src, void* dst, long n16, hipStream_t st) {
  hipLaunchKernelGGL(k_copy16, dim3(1024), dim3(256), 0, st, (const float4*)src, (float4*)dst, n16);
  return hipGetLastError();
}

hipError_t launch_assemble(const SysDev* s, int nsys, int maxM, hipStream_t st) {
  hipLaunchKernelGGL(k_assemble, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_exp_prop(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st) {
  if (prec == 0) hipLaunchKernelGGL(k_exp_prop<double>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  else hipLaunchKernelGGL(k_exp_prop<float>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_graft_init(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st) {
  if (prec == 0) hipLaunchKernelGGL(k_graft_init<double>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  else hipLaunchKernelGGL(k_graft_init<float>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_graft_density(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st) {
  if (prec == 0) hipLaunchKernelGGL(k_graft_density<double>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  else hipLaunchKernelGGL(k_graft_density<float>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_ik_mult(const SysDev* one, long M, const double2* in, double2* out, int d, hipStream_t st) {
  hipLaunchKernelGGL(k_ik_mult, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, one, in, out, d);
  return hipGetLastError();
}

hipError_t launch_sites_accum(long M, const double2* a, double2* acc, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_sites_accum, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, M, a, acc, mode);
  return hipGetLastError();
}

hipError_t launch_density(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st) {
  if (prec == 0) hipLaunchKernelGGL(k_density<double>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  else hipLaunchKernelGGL(k_density<float>, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_cavity(const SysDev* s, int nsys, int maxM, hipStream_t st) {
  hipLaunchKernelGGL(k_cavity, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_plate_grad(const SysDev* one, int nblk, const double* axis, double lq_re, double lq_im,
                             double2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_plate_grad1, dim3(nblk), dim3(256), 0, st, one, axis[0], axis[1], axis[2], lq_re, lq_im);
  hipLaunchKernelGGL(k_plate_grad2, dim3(1), dim3(256), 0, st, one, out);
  return hipGetLastError();
}

hipError_t launch_site_shape(const SysDev* one, long M, int p, int lam, const double* axis, double scale,
                             double2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_site_shape, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, one, p, lam,
                     axis ? axis[0] : 0.0, axis ? axis[1] : 0.0, axis ? axis[2] : 0.0, scale, out);
  return hipGetLastError();
}

hipError_t launch_cmul_acc(long M, const double2* a, const double2* b, double2* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_cmul_acc, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, M, a, b, acc);
  return hipGetLastError();
}

hipError_t launch_site_grad_partial(int prec, const SysDev* one, int nblk, const double2* S, const double2* g,
                                    double q_re, double q_im, double g2_min, double2* partials, hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL(k_site_grad_partial<double>, dim3(nblk), dim3(256), 0, st, one, S, g, q_re, q_im, g2_min,
                       partials);
  else
    hipLaunchKernelGGL(k_site_grad_partial<float>, dim3(nblk), dim3(256), 0, st, one, S, g, q_re, q_im, g2_min,
                       partials);
  return hipGetLastError();
}

hipError_t launch_scale_real(double* a, long n, double f, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_real, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n, f);
  return hipGetLastError();
}

hipError_t launch_noise(const SysDev* s, int nsys, int maxM, hipStream_t st) {
  hipLaunchKernelGGL(k_noise, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_density_seg(int prec, const SysDev* s, int nsys, int maxM, int lo, int n, int qd0, int first,
                              int last, hipStream_t st) {
  if (prec == 0) hipLaunchKernelGGL(k_density_seg<double>, pgrid(maxM, nsys), dim3(256), 0, st, s, lo, n, qd0, first, last);
  else hipLaunchKernelGGL(k_density_seg<float>, pgrid(maxM, nsys), dim3(256), 0, st, s, lo, n, qd0, first, last);
  return hipGetLastError();
}

hipError_t launch_ck_copy(const SysDev* s, int nsys, int maxM, int esz, int src, int src_slot, int dst, int dst_slot,
                          int dst2, int dst2_slot, hipStream_t st) {
  const long n16 = (long)maxM * esz / 16;
  hipLaunchKernelGGL(k_ck_copy, dim3((unsigned)((n16 + 255) / 256), nsys), dim3(256), 0, st, s, esz, src, src_slot,
                     dst, dst_slot, dst2, dst2_slot);
  return hipGetLastError();
}

hipError_t launch_noise_advance(const SysDev* s, int nsys, hipStream_t st) {
  hipLaunchKernelGGL(k_noise_advance, dim3((nsys + 63) / 64), dim3(64), 0, st, s, nsys);
  return hipGetLastError();
}

hipError_t launch_cl_accumulate(const SysDev* s, int nsys, int maxM, hipStream_t st) {
  hipLaunchKernelGGL(k_cl_accumulate, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_update(int scheme, const SysDev* s, int nsys, int maxM, hipStream_t st) {
  if (scheme == 1) hipLaunchKernelGGL(k_update_1s, pgrid(maxM, nsys), dim3(256), 0, st, s);
  else hipLaunchKernelGGL(k_update_euler, pgrid(maxM, nsys), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_reduce(int what, int prec, const SysDev* s, int nsys, int nblk, hipStream_t st) {
  dim3 g1(nblk, nsys), g2(1, nsys);
  switch (what) {
    case RED_Q:
      if (prec == 0) hipLaunchKernelGGL((k_reduce1<RED_Q, double>), g1, dim3(256), 0, st, s);
      else hipLaunchKernelGGL((k_reduce1<RED_Q, float>), g1, dim3(256), 0, st, s);
      hipLaunchKernelGGL(k_reduce2<RED_Q>, g2, dim3(256), 0, st, s);
      break;
    case RED_H:
      hipLaunchKernelGGL((k_reduce1<RED_H, double>), g1, dim3(256), 0, st, s);
      hipLaunchKernelGGL(k_reduce2<RED_H>, g2, dim3(256), 0, st, s);
      break;
    case RED_SHIFT:
      hipLaunchKernelGGL((k_reduce1<RED_SHIFT, double>), g1, dim3(256), 0, st, s);
      hipLaunchKernelGGL(k_reduce2<RED_SHIFT>, g2, dim3(256), 0, st, s);
      break;
    case RED_VEXP:
      hipLaunchKernelGGL((k_reduce1<RED_VEXP, double>), g1, dim3(256), 0, st, s);
      hipLaunchKernelGGL(k_reduce2<RED_VEXP>, g2, dim3(256), 0, st, s);
      break;
    case RED_QGA:
      if (prec == 0) hipLaunchKernelGGL((k_reduce1<RED_QGA, double>), g1, dim3(256), 0, st, s);
      else hipLaunchKernelGGL((k_reduce1<RED_QGA, float>), g1, dim3(256), 0, st, s);
      hipLaunchKernelGGL(k_reduce2<RED_QGA>, g2, dim3(256), 0, st, s);
      break;
    default:
      hipLaunchKernelGGL((k_reduce1<RED_RHO, double>), g1, dim3(256), 0, st, s);
      hipLaunchKernelGGL(k_reduce2<RED_RHO>, g2, dim3(256), 0, st, s);
      break;
  }
  return hipGetLastError();
}

}  // namespace scft
