/*
 * channelflow.h -- C-ABI of the MI355X-native RK3 timestep of the
 * pseudo-spectral channel-flow LES (reference: fxie2/paes-channelflow, Fortran 90).
 *
 * One context owns one GPU, one HIP stream and every device buffer of the run
 * (fields, RK registers, pressure, 3/2-grid intermediates, FFT twiddles).  The
 * host never holds device pointers: fields cross the boundary as plain
 * column-major REAL(8) host arrays (x fastest, then y, then z), exactly the
 * memory order of the reference's u(nnx,nny,iz) planes (modules.f90:21-26).
 *
 * The reference has no C interface; its hot path is a set of Fortran
 * subroutines with implicit interfaces that read and write module globals.
 * Each entry point below names the reference code it replaces.  The Fortran
 * driver (paes-channelflow_amd/fortran/les_gpu.F90) binds them with
 * ISO_C_BINDING.
 *
 * Return codes: 0 = success, < 0 = error (message via cf_last_error).
 * Threading: a context is used by one host thread at a time.  Every call is
 * synchronous on return except cf_stage/cf_step, which enqueue work on the
 * context's stream; errors raised on the device (the utau > 10 guard, a
 * non-finite field) are reported by the next cf_sync / cf_get_* call.
 */
#ifndef CHANNELFLOW_H
#define CHANNELFLOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CF_OK 0
#define CF_E_ARG (-1)         /* invalid argument or unsupported option          */
#define CF_E_HIP (-2)         /* HIP runtime error (allocation, launch, copy)     */
#define CF_E_UTAU (-3)        /* utau > 10 wall-model guard: lowerbc.f90:77-80,    */
                              /* upperbc.f90:76-79 (the reference calls lesabort) */
#define CF_E_NONFINITE (-4)   /* a field became NaN/Inf                           */
#define CF_E_ECLAMP (-5)      /* reserved: no longer returned (the comp1 clamp of  */
                              /* the expanded e, comp1.f90:109,121, is carried out) */
#define CF_E_NODEV (-6)       /* no HIP device                                     */

#define CF_PREC_FP64 0        /* every field and transform in FP64                  */
#define CF_PREC_FP32 1        /* FP32 fields/FFTs/products/SGS; FP64 pressure RHS,  */
                              /* tridiagonal solve, projection and update           */

typedef struct cf_ctx cf_ctx;

/* Run parameters: the namelist groups GRIDSIZE/PARAMS/TIMING of les.F90:30-34
 * plus the precision mode.  Derived constants follow les.F90:92-121. */
typedef struct cf_params {
  int32_t nnx, nny, nnz;   /* grid; nnx == nny in {32,64,128,256,512}, nnz >= 4 */
  double xl, yl, zl;       /* domain lengths                                    */
  double z0;               /* roughness length (wall model)                     */
  double dt;               /* fixed time step (xkmax = dz^2/dt/5, les.F90:216)   */
  int32_t imodel;          /* SGS diffusivity: 3 TKE (Moeng 84), 4 Smagorinsky   */
                           /* (complen.f90:19-24); the learned closure needs 3.  */
  int32_t isurface;        /* wall model 1..6 (lowerbc.f90:126-161): 1 Schumann, */
                           /* 2 Moeng, 3 Wei, 4 Moeng unlinearised, 5 Wei with   */
                           /* Moeng's fluctuation, 6 mean-forcing balance.  The  */
                           /* reference's uninitialised ufree is taken as 0.     */
  int32_t precision;       /* CF_PREC_FP64 or CF_PREC_FP32                       */
  int32_t device;          /* HIP device ordinal                                 */
  double alk;              /* SGS length of imodel 3; <= 0 selects dsl.  The    */
                           /* reference never assigns it (allocate.f90:61).     */
  double ugal, vgal;       /* Galilean shifts of the wall model (storage module) */
} cf_params;

/* Set-up: replaces setcombuf/allocate_arrays/initwave/fft_init
 * (les.F90:146-149; buffers.f90:1-34, allocate.f90:1-127, init.f90:4-44,
 * filter_fftw.f90:16-75).  Allocates every device buffer once. */
int cf_create(const cf_params* p, cf_ctx** out);

/* Teardown: fft_finalize + deallocate_arrays (les.F90:356-357). */
int cf_destroy(cf_ctx* ctx);

/* Load the interior levels 1..nnz of u, v, w, e (each nnx*nny*nnz doubles;
 * w at faces 1..nnz, w(nnz) is the top wall) and apply deal() to all four
 * (les.F90:185-190, filter_fftw.f90:316-329).  Replaces read_field/initvel
 * (les.F90:158-173).  RK registers and pressure are zeroed. */
int cf_set_state(cf_ctx* ctx, const double* u, const double* v, const double* w,
                 const double* e);

/* Seeded initial condition, written to the device as by cf_set_state.
 * u = u_tau/0.4 ln(z_w/z0) + amp u_tau xi, v = amp u_tau xi, w = amp u_tau xi
 * (faces 1..nnz-1; w(nnz) = 0), e = 0.01 u_tau^2, with u_tau = sqrt(0.016 zl/2),
 * z_w = min(zm, zl-zm), zm = (iz-1/2) dz (the log profile of init.f90:71-78)
 * and xi in [-1,1) from SplitMix64 keyed by (seed, field, i, j, iz).  Replaces
 * initvel (init.f90:50-114), whose random stream is compiler-specific. */
int cf_seeded_init(cf_ctx* ctx, uint64_t seed, double amp);

/* Read the interior levels 1..nnz of u, v, w, e and p (any pointer may be
 * NULL).  Replaces the per-level record of sav_field (read_write.f90:31-40). */
int cf_get_state(cf_ctx* ctx, double* u, double* v, double* w, double* e, double* p);

/* Read the RK registers ru, rv, rw, re (levels 1..nnz, pressure gradient
 * included: the values comp2 leaves behind, comp2.f90:17-31).  Any pointer
 * may be NULL. */
int cf_get_tendency(cf_ctx* ctx, double* ru, double* rv, double* rw, double* re);

/* Read the comp1 tendencies of the last stage before the pressure projection
 * (ru, rv, rw after compress and the rw mean removal, comp1.f90:161-177). */
int cf_get_comp1_tendency(cf_ctx* ctx, double* ru, double* rv, double* rw);

/* One RK3 stage istage in {1,2,3} with the fixed dt: the body of the stage
 * loop les.F90:254-285 -- lower, upper, comp1, comprhs, press, comp2.  When
 * mnout != 0 and istage == 1, the SGS stress means of compsgs3.f90:71-78 are
 * accumulated.  Enqueued on the context stream. */
int cf_stage(cf_ctx* ctx, int istage, int mnout);

/* nsteps full RK3 steps (3 stages each) without means output.  Enqueued. */
int cf_step(cf_ctx* ctx, int nsteps);

/* Change dt (and xkmax) for subsequent stages. */
int cf_set_dt(cf_ctx* ctx, double dt);

/* Wait for the stream and report any device-side error flag. */
int cf_sync(cf_ctx* ctx);

/* Mean-square 3/2-grid divergence of the last stage's input field per level
 * (divz, comp1.f90:145), nnz doubles. */
int cf_get_divz(cf_ctx* ctx, double* divz);

/* utau, utau_s, dudt, dvdt as left by sufsim/sufsim_upper (lowerbc.f90:67-72,
 * upperbc.f90:74): 4 doubles. */
int cf_get_scalars(cf_ctx* ctx, double* scalars);

/* Mean profiles (compmn.F90:1-64) into meanbuf[19*nnz] in the row order of
 * allocate.f90:89-111 (englez, engsbz, divz, uxym, vxym, wxym, uule, vvle,
 * wwle, uwle, vwle, uvle, pple, uusb, vvsb, uvsb, wwsb, uwsb, vwsb), stored
 * meanbuf[row + 19*(iz-1)] like the Fortran meanbuf(19,nnz).  Computed on
 * the device from the band spectra (Parseval); only the profiles cross PCIe. */
int cf_compmn(cf_ctx* ctx, double* meanbuf);

/* x-y energy spectra of the current state (calc_spectra.f90:139-243 for one
 * field snapshot, fcounter = 1) at the 1-based levels zlevels[nlev], each in
 * 1..nnz-1.  u and v average levels zl and zl+1 with their plane means
 * removed; w is taken at face zl.  Outputs:
 * - hats[3][nlev][nnx/2+1]: ring spectra of u, v, w, normalised as the
 *   reference does (bin 0 left unnormalised);
 * - var[3][nlev];
 * - lscales[6][nlev]: l111 l112 l221 l222 l331 l332.
 * Computed on the device from the band spectra. */
int cf_calc_spectra(cf_ctx* ctx, int32_t nlev, const int32_t* zlevels, double* hats, double* var,
                    double* lscales);

/* Allocation-free variant of cf_calc_spectra.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as cf_calc_spectra, for at most 64 levels
 * (CF_E_ARG beyond).  It calls no hipMalloc/hipFree: its scratch (ring sums,
 * normalised outputs and the level list, about 0.8 MB) is a fixed device
 * array held by each GPU until the process ends and shared by every call on
 * that GPU, so calls that overlap on one GPU race. */
int cf_calc_spectra_noalloc(cf_ctx* ctx, int32_t nlev, const int32_t* zlevels, double* hats, double* var,
                            double* lscales);

/* Time nsteps steps with HIP events on the context stream.  Writes the total
 * milliseconds and, when kernel_ms is not NULL, the average duration of each
 * kernel family of one step measured with events bracketing every launch
 * (kernel_ms[CF_NKERN]; the per-kernel pass runs after the total timing). */
/* k_wall k_col_synth k_divz_band k_rows k_col_anal | k_sgs_cols_in k_sgs_rows
 * k_sgs_cols_out (the small-grid closure: zero time otherwise) | k_press k_erows
 * k_reduce_part (the SGS means of an mnout stage: zero time otherwise).  A family
 * no kernel of which runs on the path reports 0. */
#define CF_NKERN 11
int cf_time_steps(cf_ctx* ctx, int nsteps, double* total_ms, double* kernel_ms);

/* Name of kernel family k of cf_time_steps (0 <= k < CF_NKERN). */
const char* cf_kernel_name(int k);

/* Bytes of device memory the context holds. */
int64_t cf_device_bytes(const cf_ctx* ctx);

/* Text of the last error of this context (or of cf_create when ctx is NULL). */
const char* cf_last_error(const cf_ctx* ctx);

/* ------------------------------------------------------------------------
 * Learned SGS closure (no counterpart in the reference).
 *
 * A pointwise MLP 11 -> 64 -> 64 -> 64 -> 6 (GELU, tanh form) maps the
 * resolved velocity gradients at a cell centre (ux uy uz vx vy vz wx wy wz),
 * the filter width dsl (les.F90:118) and the wall distance min(z, zl-z) to
 * the SGS stresses (t11 t22 t33 t12 t13 t23).  It replaces the stresses of
 * compsgs1 (compsgs1.f90:1-53) at the same staggered locations, with the
 * TKE closure as fallback wherever the prediction is not realizable or
 * backscatters.  BF16 weights and activations on MFMA with FP32
 * accumulation; FP32 master weights and Adam (beta 0.9/0.999, eps 1e-8).
 *
 * Inputs are standardised x = (f - in_mu)/in_sd and outputs de-standardised
 * tau = y*out_sd + out_mu.  Parameters are one flat float array in the
 * order W1[64][11], b1[64], W2[64][64], b2, W3[64][64], b3, W4[6][64],
 * b4[6] (CF_MLP_NPARAM floats).
 * ------------------------------------------------------------------------ */
typedef struct cf_mlp cf_mlp;
#define CF_MLP_NIN 11
#define CF_MLP_NOUT 6
#define CF_MLP_NPARAM 9478

/* Network object on a device: parameters, optimiser state, training samples. */
int cf_mlp_create(int32_t device, cf_mlp** out);
int cf_mlp_destroy(cf_mlp* m);

/* Flat parameters (CF_MLP_NPARAM floats) in and out. */
int cf_mlp_set_params(cf_mlp* m, const float* params);
int cf_mlp_get_params(cf_mlp* m, float* params);

/* Standardisation: in_mu, in_sd (11 each), out_mu, out_sd (6 each); sd > 0.
 * Must be set before cf_mlp_put_samples, which stores samples standardised. */
int cf_mlp_set_norm(cf_mlp* m, const float* in_mu, const float* in_sd, const float* out_mu,
                    const float* out_sd);
int cf_mlp_get_norm(cf_mlp* m, float* in_mu, float* in_sd, float* out_mu, float* out_sd);

/* Network file: "CFMLP001", int32 {11, 64, 6, 9478}, float in_mu[11],
 * in_sd[11], out_mu[6], out_sd[6], params[9478], little endian. */
int cf_mlp_save(cf_mlp* m, const char* path);
int cf_mlp_load(cf_mlp* m, const char* path);

/* Inference on the device: feats [n][11] raw -> stresses [n][6] raw. */
int cf_mlp_forward(cf_mlp* m, int64_t n, const float* feats, float* stresses);

/* Training samples on the device: room for capacity samples (discards any
 * stored), then raw features [n][11] and target stresses [n][6] at offset. */
int cf_mlp_reserve(cf_mlp* m, int64_t capacity);
int cf_mlp_put_samples(cf_mlp* m, int64_t offset, int64_t n, const float* feats,
                       const float* targets);

/* Mean squared error over the standardised outputs of the stored samples
 * offset..offset+n-1 and its gradient (CF_MLP_NPARAM floats); no update. */
int cf_mlp_loss_grad(cf_mlp* m, int64_t offset, int64_t n, float* loss, float* grad);

/* nsteps Adam steps at learning rate lr.  batch > 0: each step draws batch
 * samples with replacement from the stored ones (counter-based, from seed
 * and the step count); batch == 0: full batch in order.  losses (nsteps
 * floats, may be NULL) receives each step's loss before its update. */
int cf_mlp_train(cf_mlp* m, int32_t nsteps, int64_t batch, float lr, uint64_t seed, float* losses);
int cf_mlp_reset_optimizer(cf_mlp* m);
const char* cf_mlp_last_error(const cf_mlp* m);

/* Use network m as the SGS closure of ctx from the next stage on (its
 * current parameters are copied); NULL restores the TKE closure.  Needs
 * nnx in {128, 256, 512}. */
int cf_set_closure(cf_ctx* ctx, const cf_mlp* m);

/* Where the network is evaluated.
 * - CF_CLOSURE_GRID_DEALIASED (default): at the 3/2-grid cell centres,
 *   inside the row-tendency sweep, where the TKE stresses of compsgs1 live.
 * - CF_CLOSURE_GRID_SMALL: at the nnx x nny cell centres, the grid the
 *   training features are computed on, with the TKE fallback on the same
 *   grid.  The stress divergence is added to the compressed tendencies, and
 *   the wall-model faces stay on the 3/2 grid.  It needs 2.25x fewer network
 *   evaluations.
 * Applies from the next stage. */
#define CF_CLOSURE_GRID_DEALIASED 0
#define CF_CLOSURE_GRID_SMALL 1
int cf_set_closure_grid(cf_ctx* ctx, int32_t grid);

/* Cells where the network's stress was used, and cells evaluated, since the
 * previous call (then reset): counts[2]. */
int cf_get_closure_counts(cf_ctx* ctx, int64_t* counts);

/* Training data: box filter of the current state by fx x fy x fz fine cells
 * (fx, fy in {1,2,4,8}, fz in {1,2,4}, dividing the grid) into
 * out[17][ncz][ncy][ncx] doubles, all box means:
 * - u, v, w (cell centre);
 * - w at the coarse face K+1 (trapezoid over fz+1 fine faces, 0 at the top
 *   wall);
 * - e (cell centre);
 * - uu, vv, ww, uv, uw, vw of the cell-centred fine velocities;
 * - the fine grid's own TKE-closure stresses t11 t22 t33 t12 t13 t23 at cell
 *   centres (compsgs1.f90:41-50; face stresses averaged to the centre). */
int cf_filter_snapshot(cf_ctx* ctx, int32_t fx, int32_t fy, int32_t fz, double* out);

#ifdef __cplusplus
}
#endif

#endif /* CHANNELFLOW_H */
