/*
 * scft.h -- C ABI of the MI355X SCFT / complex-Langevin hot path.
 *
 * The reference (rar-ensemble/aligned-nanoplate-separation-and-rotation) is one
 * MPI + FFTW C++ program whose hot path communicates through process globals
 * (code/globals.h:37-167).  This header replaces those seams with a context
 * object and plain C types: interleaved complex<double> arrays (re, im) of
 * length M in the reference layout, linear index x + Nx*(y + Ny*z), x fastest
 * (code/array_utils.cpp:80-87).  Every call returns SCFT_OK (0) or a negative
 * error code; nothing calls exit().  A NULL context or array pointer returns
 * SCFT_EINVAL, with two exceptions: the set/get calls that move several
 * arrays at once (scft_set_fields, scft_get_fields, scft_get_density,
 * scft_get_smeared, scft_get_noise, scft_get_averages) skip a NULL array, and
 * scft_destroy(NULL) is a no-op.  scft_last_error() describes the last
 * failure of the calling thread.
 *
 * Threading: one host thread per context.  Calls on one context are ordered
 * on the context's HIP stream; the scft_get_* calls synchronise.
 */
#ifndef SCFT_H
#define SCFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SCFT_OK            0
#define SCFT_EINVAL       -1   /* bad argument or parameter set            */
#define SCFT_ENOMEM       -2   /* device or host allocation failed         */
#define SCFT_EDEVICE      -3   /* HIP runtime error                        */
#define SCFT_ENAN         -4   /* H or Q is NaN (code/calc_h.cpp:38-61,
                                  code/simulate.cpp:100-111)               */
#define SCFT_EUNSUPPORTED -5   /* grid size / option not built            */
#define SCFT_EIO          -6   /* file could not be read or written       */

/* ---- precision of the propagator / FFT stage --------------------------- */
#define SCFT_PREC_FP64 0       /* plain complex128 everywhere              */
#define SCFT_PREC_FP32 1       /* complex64 propagator and FFTs; fields,
                                  Q, log Q, H and residual stay FP64       */
/* Option OR-ed into `precision` of scft_create: checkpointed propagators.
 * Instead of q and q-dagger at all N contour steps (2 N arrays per system,
 * code/array_utils.cpp:509-515) the context keeps q at ceil(N/K) checkpoints
 * (K = ceil(sqrt N)) and two windows of K + 1 steps, about 3 sqrt(N) arrays:
 * each density evaluation recomputes the q chain segment by segment while
 * the q-dagger chain advances (3(N-1) steps instead of 2(N-1)).  Needs the
 * same N in all systems and 0 < Nda < N. */
#define SCFT_OPT_CHECKPOINT 0x100

/* ---- plate shapes ------------------------------------------------------ */
#define SCFT_NP_SPHERE 1       /* reference np_type 1 (not on this path)   */
#define SCFT_NP_ROD    2       /* reference np_type 2: 1/4 erfc*erfc,
                                  code/initialize.cpp:578-612              */
#define SCFT_NP_SQUARE 3       /* extension: square plate, 1/8 erfc^3      */

/*
 * Simulation parameters.  Field order follows the positional bcp.input file
 * read by read_input() (code/read_input.cpp:31-236).  chi values are stored
 * signed, exactly as written in the file; the negative-chi flags of the
 * reference (code/read_input.cpp:53-85) are derived from the sign.
 */
typedef struct scft_params {
  int dim;                         /* 2 or 3 (reference: compile-time Dim)  */
  /* main parameters, code/read_input.cpp:31-47 */
  int Nda, Ndb;
  double phiH;
  int Nah;
  double a_smear;
  double rho0;
  /* interaction parameters, code/read_input.cpp:53-87 */
  double chi_ab, chi_bc, chi_ac;
  double kappa;
  /* initial-condition flags, code/read_input.cpp:103-108 */
  int ic_flag[2];
  double ic_pre[2], ic_dir[2], ic_period[2];
  /* numerical parameters, code/read_input.cpp:114-135 */
  int Nx[3];
  double L[3];
  double lam_imag, lam_real;
  int itermax, print_freq, sample_freq, sample_wait;
  double error_tol;
  int update_scheme;               /* 0 = Euler, 1 = 1S                    */
  int do_CL;
  /* film, code/read_input.cpp:141-146 */
  int do_film;
  double wallT, wallXi, top_wall_lamA, top_wall_lamB, bot_wall_lamA,
      bot_wall_lamB;
  /* channel, code/read_input.cpp:155-158 */
  int do_channel, hor_dir, vert_dir;
  double wall_width, channel_width, channel_xi;
  /* nanoparticles, code/read_input.cpp:182-236 */
  int n_exp_nr, do_fld_np, np_type, np_chem;
  double sigma;
  int Ng, Nu;
  double np_frac;
  double L_nr, R_nr, xi_nr;
  double exp_nr_c[2][3];           /* fractional centres (0..1)            */
  double exp_nr_u[2][3];           /* unit orientation, normalised on read */
  /* extension: seed of the counter-based noise generator (CL noise) and of
   * ic_flag == -1 random starts.  The reference seeds ran2 from the clock
   * (code/main.cpp:20). */
  uint64_t seed;
} scft_params;

/* Scalars fixed at set-up (code/initialize.cpp:45-81, 249-355). */
typedef struct scft_sysinfo {
  int M;                           /* grid points                          */
  int N;                           /* Nda + Ndb                            */
  double V, Vf, dx[3];
  double V_1_exp_np, np_frac, phiD, phiG, nP, ng_per_np, nD;
  double Rg2, fD;
} scft_sysinfo;

typedef struct scft_ctx scft_ctx;

/* Parse a bcp.input file (code/read_input.cpp:7-242) for a run in `dim`
 * dimensions: each grid/box/centre/orientation line holds `dim` values and
 * any further number on the line is skipped, as fgets() does there. */
int scft_read_input(const char* path, int dim, scft_params* out);

/* Create a context holding `n_systems` systems on HIP device `device`.
 * `params` points to n_systems structs (boxes may differ).  Performs the
 * set-up of initialize_1/initialize_2 (code/initialize.cpp:45-412): smearing
 * and bond kernels, plate densities, graft sites, volume bookkeeping, Debye
 * functions and the initial fields of init_fields (code/initialize.cpp:
 * 419-565), then times the size-specialised pass variants on the context's
 * own buffers and keeps the fastest.  Environment switches read here:
 * SCFT_NO_TUNE (keep the first variant of each pass, no timing),
 * SCFT_TUNE_PICK=k (take variant k, or the last, of every pass; for tests),
 * SCFT_ONE_STREAM (run the q and q-dagger chains on one stream). */
int scft_create(const scft_params* params, int n_systems, int device,
                int precision, scft_ctx** out);
int scft_destroy(scft_ctx* ctx);
int scft_get_info(scft_ctx* ctx, int sys, scft_sysinfo* out);

/* Field exchange with host memory: w+, wAB+, wAB- (code/globals.h:76-90).
 * Replaces the .res override of read_resume_files (code/io_utils.cpp:183-240). */
int scft_set_fields(scft_ctx* ctx, int sys, const double* wpl,
                    const double* wabp, const double* wabm);
int scft_get_fields(scft_ctx* ctx, int sys, double* wpl, double* wabp,
                    double* wabm);
/* Diblock densities rhoda, rhodb in real space (code/globals.h:99-100). */
int scft_get_density(scft_ctx* ctx, int sys, double* rhoA, double* rhoB);
/* Smeared fields smwa, smwb (code/calc_poly_density.cpp:93-104). */
int scft_get_smeared(scft_ctx* ctx, int sys, double* smwa, double* smwb);
/* Plate density rho_exp_nr (code/initialize.cpp:210-239). */
int scft_get_cavity(scft_ctx* ctx, int sys, double* rho);
/* Single-chain partition function Qd of the last density evaluation. */
int scft_get_Q(scft_ctx* ctx, int sys, double* Q_re_im);

/* Named arrays of M interleaved complex values (the reference's globals,
 * code/globals.h:76-123, and the .res / .bin files of code/io_utils.cpp).
 * Settable fields: "wpl", "wabp", "wabm" and, for systems with BC/AC field
 * pairs (grafted chains, chi_BC or chi_AC nonzero), "wbcp", "wbcm", "wacp",
 * "wacm" (the .res overrides of
 * read_resume_files, code/io_utils.cpp:183-240).  Readable in addition:
 * "rhoda", "rhodb", "smwa", "smwb", "rho_exp_nr", with rod plates
 * "expl_grafts" (graft sites, code/initialize.cpp:160-205, computed for rods
 * whether or not chains are grafted) and, with grafts, "smwg", "rhog_exp"
 * (graft density) and, in FP64, "qg:j" (free-end graft propagator at
 * contour index j).  SCFT_EINVAL for a name the system does not hold. */
int scft_get_array(scft_ctx* ctx, int sys, const char* name, double* out);
int scft_set_array(scft_ctx* ctx, int sys, const char* name, const double* in);
/* Qga_exp = integ(expl_grafts log q_g[Ng-1]) of the last density evaluation
 * (code/calc_poly_density.cpp:155-160); SCFT_EINVAL without grafts. */
int scft_get_Qga(scft_ctx* ctx, int sys, double* Qga_re_im);

/* Move or rotate the plates of system `sys` in place: `centres` and `u` hold
 * 3 doubles per plate (fractional centres 0..1, orientation vectors,
 * normalised here), as the plate lines of bcp.input (code/read_input.cpp:
 * 205-236).  Recomputes the plate density on the device
 * (code/initialize.cpp:578-612), its transform exp_nrH, the volume
 * bookkeeping nD, phi_D (code/initialize.cpp:249-355) and, with grafted
 * chains, the graft sites; fields are kept. */
int scft_set_plates(scft_ctx* ctx, int sys, const double* centres,
                    const double* u);

/* Explicit derivatives of H at the current fields with respect to each
 * plate's centre (absolute coordinates, b units) and to a rotation of the
 * plate about `axis` (unit vector) through its centre.  H depends on the
 * plates through rho_np in wpl_part (code/calc_h.cpp:71-72) and through nD
 * in -nD log Qd (code/initialize.cpp:351):
 *   dH/dg = sum_r d rho_np/dg (i rho0 w+ + (rho0 / N) log Qd) dV.
 * At a saddle point (dH/dw = 0) these are the total derivatives: the force
 * and torque on the plates.  out[8 p + 2 k + {0,1}] = Re, Im of dH/dc_k
 * (k = 0..2) and of dH/dtheta (k = 3), p = 0..1.  Needs a density
 * evaluation (Qd) of the current fields.  With grafted chains H also depends
 * on the plates through the graft sites (-n ng_per_np Re Qga,
 * code/calc_h.cpp:31-35); that term is included (real parts), as the exact
 * derivative of the discrete graft sites of code/initialize.cpp:160-196. */
int scft_grad_plates(scft_ctx* ctx, int sys, const double* axis, double* out);

/* Complex-Langevin noise of the last update, in k-space (the reference's
 * etap / etam globals, code/globals.h:111, filled by generate_1s_noise,
 * code/1s_update.cpp:356-368): the w+ noise, the AB+ noise and the AB- noise.
 * SCFT_EINVAL if the system has do_CL = 0. */
int scft_get_noise(scft_ctx* ctx, int sys, double* eta_wpl, double* eta_abp,
                   double* eta_abm);

/* Complex-Langevin density averages, accumulate_all_averages
 * (code/array_utils.cpp:13-35) as sampled by the main loop: scft_iterate adds
 * rhoA and rhoB after the update of every iteration it with
 * it >= sample_wait and it % sample_freq == 0 (code/simulate.cpp:91-92),
 * it counting updates from 0 at create.  Returns the sums divided by the
 * sample count as write_avg_data_bin does (code/io_utils.cpp:368-369; sums,
 * i.e. zeros, while the count is 0) and the count in *n_samples.  Any output
 * pointer may be NULL.  SCFT_EINVAL if the system has do_CL = 0. */
int scft_get_averages(scft_ctx* ctx, int sys, double* avg_rhoA, double* avg_rhoB,
                      double* n_samples);

/* FFT seam, fft_fwd_wrapper / fft_bck_wrapper (code/fft_wrappers.cpp:4-67):
 * dir = -1 forward, scaled by 1/M; dir = +1 backward, unscaled.  `inout` is a
 * host array of M interleaved complex values of system `sys`'s grid. */
int scft_fft(scft_ctx* ctx, int sys, double* inout, int dir);

/* Stage calls on every system of the context. */
/* calc_poly_density() (code/calc_poly_density.cpp:37-122, 167-191): field
 * assembly, smearing, diblock propagators (code/diblock_discrete.cpp:4-95),
 * Qd and the q.q-dagger density sum. */
int scft_calc_poly_density(scft_ctx* ctx);
/* update_1s() (code/1s_update.cpp:7-352), ending in calc_poly_density(). */
int scft_update_1s(scft_ctx* ctx);
/* update_Euler() (code/Euler_update.cpp:5-213), ending in calc_poly_density(). */
int scft_update_euler(scft_ctx* ctx);
/* calc_H() (code/calc_h.cpp:9-64); H_re_im receives 2*n_systems doubles.
 * Returns SCFT_ENAN if any H is NaN. */
int scft_calc_H(scft_ctx* ctx, double* H_re_im);

/* The loop body of simulate() (code/simulate.cpp:85-157) for n_iter
 * iterations of the update selected by update_scheme, with the
 * complex-Langevin averages (scft_get_averages) but without H or output.
 * Captured into a HIP graph after the first call. */
int scft_iterate(scft_ctx* ctx, int n_iter);

/* Wait for all work queued on the context's device (hipDeviceSynchronize). */
int scft_synchronize(scft_ctx* ctx);

/* Measurement hook (no counterpart in the reference; used by the benchmark):
 * times `reps` back-to-back launches of one kernel of the hot path with HIP
 * events on the context's stream and returns the average duration and the
 * algorithmic HBM bytes one launch moves (each array read or written once).
 *   which = 0: propagator xy-plane pass (one contour step of the q chains
 *              of all systems; of both chains where they share a launch,
 *              i.e. when the environment variable SCFT_ONE_STREAM is set)
 *   which = 1: propagator z-pencil pass with exp(-w) applied on store
 *   which = 2: q.q-dagger density sum
 *   which = 3: one full 1S iteration (update + densities), in ms, bytes per
 *              the reference dataflow model
 *   which = 4: a plain copy of 2*M propagator values of system 0 (from its
 *              contour slices 0-1 to 2-3): calibration of byte counters
 * The propagator passes recompute contour step 1 from step 0, so repeating
 * them leaves the state unchanged. */
int scft_time_kernel(scft_ctx* ctx, int which, int reps, double* avg_ms,
                     double* bytes);

/* Text description of the launch plan (kernels, sizes, the variant the
 * autotuner chose for each size-specialised pass); writes at most len bytes
 * including the terminating zero.  No counterpart in the reference. */
int scft_describe(scft_ctx* ctx, char* buf, int len);

/* Text of the last error on this thread. */
const char* scft_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SCFT_H */
