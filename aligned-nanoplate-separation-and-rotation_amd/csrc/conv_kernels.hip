// Size-specialised passes: convolution IDFT[g(k) DFT[.]] (propagator,
// smearing) and forward / backward DFTs (field stage) along x and y of one
// xy plane (k_plane_op) and along z of a tile of z pencils (k_z_op), with
// register-resident line transforms
// (reg_fft.hpp).  LDS holds the tile between the coalesced global accesses
// and the lane-group layout, and the twiddle / multiplier tables.
//
// Each size has several variants (lanes per line G, tile width); the host
// autotuner times them on the real launch at context creation and keeps the
// fastest.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "reg_fft.hpp"
#include "scft_dev.hpp"

namespace scft {

constexpr int round64(int v) { return (v + 63) / 64 * 64; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int NX, int NY, int GX, int GY>
struct PlaneCfg {
  static constexpr int T = round64(cmax(NY * GX, NX * GY));
  static_assert(T <= 1024, "plane too large for one workgroup");
  static_assert(NX % GX == 0 && NY % GY == 0, "lane split");
};
// LDS row stride in complex elements: 32 bytes of padding per row
template <typename R, int NX>
constexpr int plane_px() { return NX + (sizeof(R) == 4 ? 4 : 2); }

template <typename R, int NX, int NY, int GX, int GY>
constexpr size_t plane_lds_bytes() {
  return sizeof(cpx<R>) * (NY * plane_px<R, NX>() + NX + NY) + sizeof(R) * (NX + NY);
}

// One xy plane per block: lines along x, then along y, each transformed by
// line_op<MODE> (convolution IDFT[g DFT[.]], forward or backward DFT) in
// lane groups; a coalesced epilogue applies the optional pointwise
// multiplier (exp(-w) on the last pass of a propagator step) and the scale.
template <typename R, int MODE, int NX, int NY, int GX, int GY>
__global__ __launch_bounds__((PlaneCfg<NX, NY, GX, GY>::T)) void k_plane_op(
    const PassDesc* __restrict__ descs, int step) {
  constexpr int T = PlaneCfg<NX, NY, GX, GY>::T, PX = plane_px<R, NX>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* twx = lds + NY * PX;
  cpx<R>* twy = twx + NX;
  R* gx = reinterpret_cast<R*>(twy + NY);
  R* gy = gx + NX;
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  const int z = blockIdx.x;
  if (z >= d.nz) return;
  const int tid = threadIdx.x;
  const long poff = (long)z * (NX * NY);
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step + poff;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step + poff;
  if (tid < NX) {
    twx[tid] = reinterpret_cast<const cpx<R>*>(d.tw0)[tid];
    if (MODE == 0) gx[tid] = reinterpret_cast<const R*>(d.g0)[tid];
  }
  if (tid < NY) {
    twy[tid] = reinterpret_cast<const cpx<R>*>(d.tw1)[tid];
    if (MODE == 0) gy[tid] = reinterpret_cast<const R*>(d.g1)[tid];
  }
#pragma unroll 4
  for (int i = tid; i < NX * NY; i += T) {
    const int y = i / NX, x = i - y * NX;
    lds[y * PX + x] = in[i];
  }
  __syncthreads();
  {
    const int grp = tid / GX, t = tid - grp * GX;
    for (int row = grp; row < NY; row += T / GX) {
      cpx<R>* line = lds + row * PX;
      line_op<MODE, R, NX, GX>(t, twx, gx, [line](int i) { return line + i; });
    }
  }
  __syncthreads();
  if (NY > 1) {
    const int grp = tid / GY, t = tid - grp * GY;
    for (int col = grp; col < NX; col += T / GY) {
      cpx<R>* line = lds + col;
      line_op<MODE, R, NY, GY>(t, twy, gy, [line](int i) { return line + i * PX; });
    }
    __syncthreads();
  }
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp) + poff;
    scale = R(d.scale);
  }
#pragma unroll 4
  for (int i = tid; i < NX * NY; i += T) {
    const int y = i / NX, x = i - y * NX;
    cpx<R> o = lds[y * PX + x];
    if (post) o = cmul(o, post[i]);
    if (scale != R(1)) o = cscale(o, scale);
    out[i] = o;
  }
}

// The same plane operation without the staging copies: x lines are loaded
// from global memory straight into the lane groups, their results go to LDS
// once, and the y lines are stored from registers to global memory with the
// multiplier and scale applied (lanes of consecutive groups hold consecutive
// x, so each store covers whole 128-byte rows).  In place (in == out) is
// safe: a block reads its whole plane before the barrier.
template <typename R, int MODE, int NX, int NY, int GX, int GY>
__global__ __launch_bounds__((PlaneCfg<NX, NY, GX, GY>::T)) void k_plane_direct(
    const PassDesc* __restrict__ descs, int step) {
  static_assert(NY > 1, "direct plane pass needs a y axis");
  constexpr int T = PlaneCfg<NX, NY, GX, GY>::T, PX = plane_px<R, NX>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* twx = lds + NY * PX;
  cpx<R>* twy = twx + NX;
  R* gx = reinterpret_cast<R*>(twy + NY);
  R* gy = gx + NX;
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  const int z = blockIdx.x;
  if (z >= d.nz) return;
  const int tid = threadIdx.x;
  const long poff = (long)z * (NX * NY);
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step + poff;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step + poff;
  if (tid < NX) {
    twx[tid] = reinterpret_cast<const cpx<R>*>(d.tw0)[tid];
    if (MODE == 0) gx[tid] = reinterpret_cast<const R*>(d.g0)[tid];
  }
  if (tid < NY) {
    twy[tid] = reinterpret_cast<const cpx<R>*>(d.tw1)[tid];
    if (MODE == 0) gy[tid] = reinterpret_cast<const R*>(d.g1)[tid];
  }
  __syncthreads();
  constexpr int SZ = (int)sizeof(cpx<R>);
  constexpr unsigned PB = (unsigned)(NX * NY * SZ);
  const __amdgpu_buffer_rsrc_t rin = buf_rsrc(in, PB);
  {
    const int grp = tid / GX, t = tid - grp * GX;
    for (int row = grp; row < NY; row += T / GX) {
      cpx<R>* dst = lds + row * PX;
      const int vo = row * NX * SZ;
      line_op_io<MODE, R, NX, GX>(
          t, twx, gx, [rin, vo](int b, int o) { return buf_ld<R>(rin, vo + b * SZ, o * SZ); },
          [dst](int b, int o, cpx<R> v) { dst[b + o] = v; });
    }
  }
  __syncthreads();
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp) + poff;
    scale = R(d.scale);
  }
  const __amdgpu_buffer_rsrc_t rout = buf_rsrc(out, PB);
  const __amdgpu_buffer_rsrc_t rpost = buf_rsrc(post ? post : out, PB);
  {
    const int grp = tid / GY, t = tid - grp * GY;
    for (int col = grp; col < NX; col += T / GY) {
      const cpx<R>* src = lds + col;
      line_op_io<MODE, R, NY, GY>(
          t, twy, gy, [src](int b, int o) { return src[(b + o) * PX]; },
          [rout, rpost, post, scale, col](int b, int o, cpx<R> v) {
            const int vo = (col + b * NX) * SZ, so = o * NX * SZ;
            if (post) v = cmul(v, buf_ld<R>(rpost, vo, so));
            if (scale != R(1)) v = cscale(v, scale);
            buf_st(v, rout, vo, so);
          });
    }
  }
}

template <int NZ, int G, int TL>
struct ZCfg {
  static constexpr int T = TL * G;
  static constexpr int LP = TL + 1;  // LDS row stride (complex elements)
  static_assert(NZ % G == 0 && T <= 1024, "z tile");
};

template <typename R, int NZ, int G, int TL>
constexpr size_t z_lds_bytes() {
  return sizeof(cpx<R>) * (NZ * ZCfg<NZ, G, TL>::LP + NZ) + sizeof(R) * NZ;
}

// A tile of TL consecutive z pencils (line l = x + nx y, stride nx ny):
// coalesced load into LDS, line_op<MODE> along z in lane groups, then a
// coalesced epilogue applying the optional multiplier and the scale.
template <typename R, int MODE, int NZ, int G, int TL>
__global__ __launch_bounds__((ZCfg<NZ, G, TL>::T)) void k_z_op(const PassDesc* __restrict__ descs,
                                                               int step) {
  using C = ZCfg<NZ, G, TL>;
  constexpr int T = C::T, LP = C::LP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* tw = lds + NZ * LP;
  R* g = reinterpret_cast<R*>(tw + NZ);
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  const int tid = threadIdx.x;
  const int nlines = d.nx * d.ny;
  const long S = nlines;
  const int l0 = blockIdx.x * TL;
  if (l0 >= nlines) return;
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step;
  for (int i = tid; i < NZ; i += T) {
    tw[i] = reinterpret_cast<const cpx<R>*>(d.tw0)[i];
    if (MODE == 0) g[i] = reinterpret_cast<const R*>(d.g0)[i];
  }
#pragma unroll 4
  for (int i = tid; i < NZ * TL; i += T) {
    const int zz = i / TL, l = i - zz * TL;
    lds[zz * LP + l] = (l0 + l < nlines) ? in[(long)l0 + l + zz * S] : mk(R(0), R(0));
  }
  __syncthreads();
  {
    const int grp = tid / G, t = tid - grp * G;
    cpx<R>* line = lds + grp;
    line_op<MODE, R, NZ, G>(t, tw, g, [line](int i) { return line + i * LP; });
  }
  __syncthreads();
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp);
    scale = R(d.scale);
  }
#pragma unroll 4
  for (int i = tid; i < NZ * TL; i += T) {
    const int zz = i / TL, l = i - zz * TL;
    if (l0 + l >= nlines) continue;
    const long gi = (long)l0 + l + zz * S;
    cpx<R> o = lds[zz * LP + l];
    if (post) o = cmul(o, post[gi]);
    if (scale != R(1)) o = cscale(o, scale);
    out[gi] = o;
  }
}

// z pencils without LDS staging: lane group g of a block holds line l0 + g,
// loaded from and stored to global memory directly.  With G lanes per line
// a wavefront holds 64 / G consecutive lines, so for G <= 8 every access
// instruction covers runs of 64 / G consecutive points (128 bytes at G = 4
// in FP32).  Only the twiddle and multiplier tables sit in LDS.
template <typename R, int MODE, int NZ, int G, int TL>
__global__ __launch_bounds__((ZCfg<NZ, G, TL>::T)) void k_z_direct(const PassDesc* __restrict__ descs,
                                                                   int step) {
  using C = ZCfg<NZ, G, TL>;
  constexpr int T = C::T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* tw = reinterpret_cast<cpx<R>*>(smem_raw);
  R* g = reinterpret_cast<R*>(tw + NZ);
  const PassDesc& d = descs[blockIdx.y];
  if (d.steps_active && step_contour(step) >= d.steps_active) return;
  const int tid = threadIdx.x;
  const int nlines = d.nx * d.ny;
  const long S = nlines;
  const int l0 = blockIdx.x * TL;
  if (l0 >= nlines) return;
  for (int i = tid; i < NZ; i += T) {
    tw[i] = reinterpret_cast<const cpx<R>*>(d.tw0)[i];
    if (MODE == 0) g[i] = reinterpret_cast<const R*>(d.g0)[i];
  }
  __syncthreads();
  const int grp = tid / G, t = tid - grp * G;
  const int l = l0 + grp;
  if (l >= nlines) return;
  constexpr int SZ = (int)sizeof(cpx<R>);
  const unsigned AB = (unsigned)(S * NZ * SZ);
  const cpx<R>* in = reinterpret_cast<const cpx<R>*>(d.in) + (long)(step_slice(step) + d.in_soff) * d.in_step;
  cpx<R>* out = reinterpret_cast<cpx<R>*>(d.out) + (long)(step_slice(step) + d.out_soff) * d.out_step;
  const cpx<R>* post = nullptr;
  R scale = R(1);
  if (d.last) {
    const void* pp = (step_contour(step) < d.post_switch) ? d.post_a : d.post_b;
    if (pp) post = reinterpret_cast<const cpx<R>*>(pp);
    scale = R(d.scale);
  }
  const __amdgpu_buffer_rsrc_t rin = buf_rsrc(in, AB), rout = buf_rsrc(out, AB);
  const __amdgpu_buffer_rsrc_t rpost = buf_rsrc(post ? post : out, AB);
  const int SB = (int)S * SZ;  // bytes between z neighbours
  line_op_io<MODE, R, NZ, G>(
      t, tw, g, [rin, l, SB](int b, int o) { return buf_ld<R>(rin, l * SZ + b * SB, o * SB); },
      [rout, rpost, post, scale, l, SB](int b, int o, cpx<R> v) {
        const int vo = l * SZ + b * SB, so = o * SB;
        if (post) v = cmul(v, buf_ld<R>(rpost, vo, so));
        if (scale != R(1)) v = cscale(v, scale);
        buf_st(v, rout, vo, so);
      });
}

// ---------------------------------------------------------------------------
// variants: (NX, NY, GX, GY) for planes, (NZ, G, TL) for z tiles
// ---------------------------------------------------------------------------
// plane variant (NX, NY, GX, GY, D): D = 1 for k_plane_direct
#define SCFT_PLANE_VARIANTS(X)                                                             \
  X(8, 8, 4, 4, 0) X(16, 16, 4, 4, 0) X(24, 16, 4, 4, 0) X(32, 32, 4, 4, 0)                \
  X(32, 32, 8, 8, 0) X(56, 64, 4, 4, 0) X(56, 64, 8, 8, 0) X(64, 64, 4, 4, 0)              \
  X(64, 64, 8, 8, 0) X(64, 64, 16, 16, 0) X(72, 64, 4, 4, 0) X(72, 64, 8, 8, 0)            \
  X(96, 96, 4, 4, 0) X(96, 96, 8, 8, 0)                                                    \
  X(32, 32, 4, 4, 1) X(32, 32, 8, 8, 1) X(56, 64, 4, 4, 1) X(56, 64, 8, 8, 1)              \
  X(64, 64, 4, 4, 1) X(64, 64, 8, 8, 1) X(72, 64, 4, 4, 1) X(72, 64, 8, 8, 1)              \
  X(96, 96, 4, 4, 1) X(96, 96, 8, 8, 1)
// z variant (NZ, G, TL, D): D = 1 for k_z_direct
#define SCFT_Z_VARIANTS(X)                                                                 \
  X(16, 4, 32, 0) X(32, 4, 32, 0) X(32, 8, 32, 0) X(64, 8, 32, 0) X(64, 8, 16, 0)          \
  X(128, 8, 32, 0) X(128, 16, 32, 0) X(128, 8, 16, 0) X(128, 16, 16, 0) X(192, 8, 32, 0)   \
  X(192, 16, 32, 0) X(192, 16, 16, 0)                                                      \
  X(64, 4, 16, 1) X(64, 4, 64, 1) X(64, 8, 32, 1) X(128, 4, 16, 1) X(128, 4, 64, 1)        \
  X(128, 8, 16, 1) X(128, 8, 32, 1) X(192, 8, 16, 1) X(192, 8, 32, 1)

template <typename R, int MODE, int NX, int NY, int GX, int GY, int D>
static hipError_t launch_plane_m(const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  using C = PlaneCfg<NX, NY, GX, GY>;
  constexpr size_t lds = plane_lds_bytes<R, NX, NY, GX, GY>();
  static bool attr[64] = {};
  if constexpr (D) {
    raise_lds_limit((const void*)k_plane_direct<R, MODE, NX, NY, GX, GY>, lds, attr);
    hipLaunchKernelGGL((k_plane_direct<R, MODE, NX, NY, GX, GY>), dim3(gridx, njobs), dim3(C::T), lds, st, d,
                       step);
  } else {
    raise_lds_limit((const void*)k_plane_op<R, MODE, NX, NY, GX, GY>, lds, attr);
    hipLaunchKernelGGL((k_plane_op<R, MODE, NX, NY, GX, GY>), dim3(gridx, njobs), dim3(C::T), lds, st, d,
                       step);
  }
  return hipGetLastError();
}

template <typename R, int NX, int NY, int GX, int GY, int D>
static hipError_t launch_plane_t(int mode, const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  if (mode == MODE_CONV) return launch_plane_m<R, MODE_CONV, NX, NY, GX, GY, D>(d, njobs, gridx, step, st);
  if (mode == MODE_FWD) return launch_plane_m<R, MODE_FWD, NX, NY, GX, GY, D>(d, njobs, gridx, step, st);
  return launch_plane_m<R, MODE_INV, NX, NY, GX, GY, D>(d, njobs, gridx, step, st);
}

template <typename R, int NZ, int G, int TL, int D>
constexpr size_t z_variant_lds() {
  return D ? sizeof(cpx<R>) * NZ + sizeof(R) * NZ : z_lds_bytes<R, NZ, G, TL>();
}

template <typename R, int MODE, int NZ, int G, int TL, int D>
static hipError_t launch_z_m(const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  using C = ZCfg<NZ, G, TL>;
  constexpr size_t lds = z_variant_lds<R, NZ, G, TL, D>();
  static bool attr[64] = {};
  if constexpr (D) {
    hipLaunchKernelGGL((k_z_direct<R, MODE, NZ, G, TL>), dim3(gridx, njobs), dim3(C::T), lds, st, d, step);
  } else {
    raise_lds_limit((const void*)k_z_op<R, MODE, NZ, G, TL>, lds, attr);
    hipLaunchKernelGGL((k_z_op<R, MODE, NZ, G, TL>), dim3(gridx, njobs), dim3(C::T), lds, st, d, step);
  }
  return hipGetLastError();
}

template <typename R, int NZ, int G, int TL, int D>
static hipError_t launch_z_t(int mode, const PassDesc* d, int njobs, int gridx, int step, hipStream_t st) {
  if (mode == MODE_CONV) return launch_z_m<R, MODE_CONV, NZ, G, TL, D>(d, njobs, gridx, step, st);
  if (mode == MODE_FWD) return launch_z_m<R, MODE_FWD, NZ, G, TL, D>(d, njobs, gridx, step, st);
  return launch_z_m<R, MODE_INV, NZ, G, TL, D>(d, njobs, gridx, step, st);
}

int fast_plane_variants(int prec, int nx, int ny, int* ids, int max_ids) {
  int n = 0, id = 0;
  // lane groups holding more than 256 bytes of a line run out of registers
#define X(a, b, ga, gb, dd)                                                                         \
  if (nx == a && ny == b && n < max_ids && ((a / ga) * (prec == 0 ? 16 : 8)) <= 256 &&            \
      (prec == 0 ? plane_lds_bytes<double, a, b, ga, gb>() : plane_lds_bytes<float, a, b, ga, gb>()) <= \
          160 * 1024)                                                                               \
    ids[n++] = id;                                                                                  \
  id++;
  SCFT_PLANE_VARIANTS(X)
#undef X
  return n;
}

int fast_z_variants(int prec, int nz, int* ids, int* tiles, int max_ids) {
  int n = 0, id = 0;
#define X(a, g, tl, dd)                                                                             \
  if (nz == a && n < max_ids && ((a / g) * (prec == 0 ? 16 : 8)) <= 256 &&                        \
      (prec == 0 ? z_variant_lds<double, a, g, tl, dd>() : z_variant_lds<float, a, g, tl, dd>()) <=   \
          160 * 1024) {                                                                             \
    tiles[n] = tl;                                                                                  \
    ids[n++] = id;                                                                                  \
  }                                                                                                 \
  id++;
  SCFT_Z_VARIANTS(X)
#undef X
  return n;
}

hipError_t launch_fast_plane(int prec, int mode, int variant, const PassDesc* d, int njobs, int gridx,
                             int step, hipStream_t st) {
  int id = 0;
#define X(a, b, ga, gb, dd)                                                                        \
  if (variant == id)                                                                               \
    return prec == 0 ? launch_plane_t<double, a, b, ga, gb, dd>(mode, d, njobs, gridx, step, st)   \
                     : launch_plane_t<float, a, b, ga, gb, dd>(mode, d, njobs, gridx, step, st);   \
  id++;
  SCFT_PLANE_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_fast_z(int prec, int mode, int variant, const PassDesc* d, int njobs, int gridx, int step,
                         hipStream_t st) {
  int id = 0;
#define X(a, g, tl, dd)                                                                      \
  if (variant == id)                                                                         \
    return prec == 0 ? launch_z_t<double, a, g, tl, dd>(mode, d, njobs, gridx, step, st)     \
                     : launch_z_t<float, a, g, tl, dd>(mode, d, njobs, gridx, step, st);     \
  id++;
  SCFT_Z_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace scft
