// Launch wrappers of scft_kernels.hip (internal to libscft_amd.so).
#pragma once

#include <hip/hip_runtime.h>

namespace scft {

struct PassDesc;
struct SysDev;

// Dynamic LDS above 64 KiB needs hipFuncAttributeMaxDynamicSharedMemorySize
// on the kernel, and the attribute belongs to the current device's copy of
// it.  `done` holds one flag per device for one kernel instantiation
// (devices 0..63; others set the attribute on every launch).
inline void raise_lds_limit(const void* kernel, size_t bytes, bool* done) {
  if (bytes <= 64 * 1024) return;
  int dev = -1;
  (void)hipGetDevice(&dev);
  const bool cached = dev >= 0 && dev < 64;
  if (cached && done[dev]) return;
  (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (cached) done[dev] = true;
}

// prec: 0 = complex128, 1 = complex64.  mode: MODE_CONV / MODE_FWD / MODE_INV.
// plane: k_plane (x and y of one xy plane per block) or k_line (one axis).
hipError_t launch_pass(int prec, int mode, bool plane, const PassDesc* d, int njobs, int gridx,
                       int tb, size_t lds, int step, hipStream_t st);
// size-specialised convolution passes (conv_kernels.hip): variant ids of a
// size, then launches by id
int fast_plane_variants(int prec, int nx, int ny, int* ids, int max_ids);
int fast_z_variants(int prec, int nz, int* ids, int* tiles, int max_ids);
hipError_t launch_fast_plane(int prec, int mode, int variant, const PassDesc* d, int njobs, int gridx,
                             int step, hipStream_t st);
hipError_t launch_fast_z(int prec, int mode, int variant, const PassDesc* d, int njobs, int gridx, int step,
                         hipStream_t st);
// line passes for the odd published lengths (odd_kernels.hip): variant ids
// and their lines per block for a length and axis, then launches by id
int odd_line_variants(int prec, int n, int axis, int* ids, int* cpb, int max_ids);
hipError_t launch_odd_line(int prec, int mode, int variant, const PassDesc* d, int njobs, int gridx, int step,
                           hipStream_t st);
hipError_t launch_copy16(const void* src, void* dst, long n16, hipStream_t st);
hipError_t launch_assemble(const SysDev* s, int nsys, int maxM, hipStream_t st);
hipError_t launch_exp_prop(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st);
hipError_t launch_density(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st);
// chains grafted to the plates: complementary start, density; graft sites
hipError_t launch_graft_init(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st);
hipError_t launch_graft_density(int prec, const SysDev* s, int nsys, int maxM, hipStream_t st);
hipError_t launch_ik_mult(const SysDev* one, long M, const double2* in, double2* out, int d, hipStream_t st);
hipError_t launch_sites_accum(long M, const double2* a, double2* acc, int mode, hipStream_t st);
// graft-site derivatives: inflated shape / its plate derivatives, acc += a b,
// block partial sums of (S / g)(L - Qga)
hipError_t launch_site_shape(const SysDev* one, long M, int p, int lam, const double* axis, double scale,
                             double2* out, hipStream_t st);
hipError_t launch_cmul_acc(long M, const double2* a, const double2* b, double2* acc, hipStream_t st);
hipError_t launch_site_grad_partial(int prec, const SysDev* one, int nblk, const double2* S, const double2* g,
                                    double q_re, double q_im, double g2_min, double2* partials, hipStream_t st);
hipError_t launch_cavity(const SysDev* s, int nsys, int maxM, hipStream_t st);
hipError_t launch_plate_grad(const SysDev* one, int nblk, const double* axis, double lq_re, double lq_im,
                             double2* out, hipStream_t st);
hipError_t launch_scale_real(double* a, long n, double f, hipStream_t st);
hipError_t launch_noise(const SysDev* s, int nsys, int maxM, hipStream_t st);
hipError_t launch_noise_advance(const SysDev* s, int nsys, hipStream_t st);
hipError_t launch_cl_accumulate(const SysDev* s, int nsys, int maxM, hipStream_t st);
// checkpointed propagators: density of one segment, slice copies
hipError_t launch_density_seg(int prec, const SysDev* s, int nsys, int maxM, int lo, int n, int qd0, int first,
                              int last, hipStream_t st);
hipError_t launch_ck_copy(const SysDev* s, int nsys, int maxM, int esz, int src, int src_slot, int dst, int dst_slot,
                          int dst2, int dst2_slot, hipStream_t st);
hipError_t launch_update(int scheme, const SysDev* s, int nsys, int maxM, hipStream_t st);
hipError_t launch_reduce(int what, int prec, const SysDev* s, int nsys, int nblk, hipStream_t st);

}  // namespace scft
