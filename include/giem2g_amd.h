/*
 * giem2g_amd.h -- C ABI of the MI355X-native GIEM2G integral-equation solve.
 *
 * Library: giem2g_amd/_lib/libgiem2g_amd.so (hipcc --offload-arch=gfx950).
 * One process, one GPU, no MPI split.  Complex numbers are interleaved
 * (re, im) doubles ("double _Complex" / Fortran COMPLEX(C_DOUBLE_COMPLEX)).
 *
 * Field vectors are Fortran-ordered (Nx, Ny, Nz, 3): element
 * v[x + Nx*(y + Ny*(z + Nz*c))].  The component slot order defaults to
 * (x, y, z) as in the reference CLI build; giem2g_set_component_order()
 * selects the (y, x, z) order of the reference shared library, which is
 * built with -D ExtrEMe (const_module.F90:72-86, Makefile:87).
 *
 * Section 1 mirrors the reference C ABI exactly (giem2g_interfaces.F90:
 * 33-58 for the structs, 76-254 for the functions).  Section 2 adds
 * status-returning, device-pointer and solver entry points that have no
 * reference counterpart; each says so.
 */
#ifndef GIEM2G_AMD_H
#define GIEM2G_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================== */
/* 1. Reference-compatible ABI                                           */
/* ===================================================================== */

/* GIEM2G_BACKGROUND_TYPE (giem2g_interfaces.F90:33-37): csigb points to
 * Nl+1 complex conductivities, [0] = air, displacement currents already
 * included (sigma - i*omega*eps0); thickness points to Nl-1 doubles. */
typedef struct {
  int Nl;
  void *csigb;
  void *thickness;
} GIEM2G_BACKGROUND_TYPE;

/* GIEM2G_ANOMALY_TYPE (giem2g_interfaces.F90:39-44): z -> Nz+1 layer
 * borders (m, positive down), dz -> Nz thicknesses. */
typedef struct {
  int Nx, Ny, Nz;
  double dx, dy;
  void *z;
  void *dz;
} GIEM2G_ANOMALY_TYPE;

/* GIEM2G_DATA_TYPE (giem2g_interfaces.F90:46-58). */
typedef struct {
  long long tensor_size;
  long long ie_kernel_buffer_length;
  long long fft_buffers_length;
  long long comm64;
  void *giem2g_tensor;
  void *dz;
  void *sqsigb;
  void *csigb;
  void *kernel_buffer;
  void *fft_buffer_in;
  void *fft_buffer_out;
} GIEM2G_DATA_TYPE;

/* Replaces GIEM2G_CALC_DATA_SIZES (giem2g_interfaces.F90:76-100).  Creates
 * the operator object and returns it in data->giem2g_tensor; reports
 * ie_kernel_buffer_length = Nx*Ny*2*Nz*(2*Nz+1) and fft_buffers_length =
 * 4*Nx*Ny*3*Nz complex elements (the reference's sizes for one process).
 * comm64 is ignored (one process). */
void giem2g_calc_data_sizes(GIEM2G_ANOMALY_TYPE anomaly, GIEM2G_DATA_TYPE *data);

/* Replaces GIEM2G_PREPARE_IE_KERNEL (giem2g_interfaces.F90:102-141).
 * Allocates the device tensor and FFT buffers.  The caller's
 * kernel_buffer / fft_buffer_in / fft_buffer_out are accepted and not used:
 * the tensor and spectra live in HBM (they may be NULL). */
void giem2g_prepare_ie_kernel(GIEM2G_ANOMALY_TYPE anomaly, GIEM2G_DATA_TYPE data);

/* Replaces GIEM2G_CALC_INTEGRAL_GREEN_TENSOR (giem2g_interfaces.F90:144-174):
 * FP64 Hankel-filter tensor build on the GPU for angular frequency omega.
 * As in the reference, the anomaly passed here sets the geometry of the
 * build (dx, dy and the layer borders z, from which each layer's background
 * layer Lnumber is derived, giem2g_interfaces.F90:158-166); its Nx, Ny, Nz
 * must be those of the prepared operator. */
void giem2g_calc_ie_kernel(void *ie_op, GIEM2G_BACKGROUND_TYPE bkg,
                           GIEM2G_ANOMALY_TYPE anomaly, double omega);

/* Replaces GIEM2G_CALC_FFT_OF_GREEN_TENSOR (giem2g_interfaces.F90:176-188). */
void giem2g_calc_fft_of_ie_kernel(void *ie_op);

/* Replaces GIEM2G_SET_ANOMALY_CONDUCTIVITY (giem2g_interfaces.F90:190-201).
 * csiga: complex (Nx,Ny,Nz) host array, aliased as in the reference (199):
 * the library keeps the pointer and every giem2g_apply_ie_operator reads the
 * array's current values, so it must stay alive while the operator is
 * applied.  The values are also uploaded at this call for the extension
 * entry points of section 2, which use the last uploaded values;
 * giem2g_op_set_sigma(_dev) replaces the alias by a copy. */
void giem2g_set_anomaly_conductivity(void *ie_op, void *csiga);

/* Replaces GIEM2G_APPLY_INTEGRAL_OPERATOR (giem2g_interfaces.F90:203-232):
 * out = A*in, both complex (Nx,Ny,Nz,3) HOST arrays (copied over PCIe),
 * with the current values of the aliased conductivity. */
void giem2g_apply_ie_operator(void *ie_op, void *in, void *out);

/* Replaces GIEM2G_SET_LOGGER (giem2g_interfaces.F90:234-238). */
void giem2g_set_logger(void (*logger)(const char *));

/* Replaces GIEM2G_PRINT_STATS (giem2g_interfaces.F90:240-247): logs the
 * operator counters in the reference's split (time for multiplications and
 * for dot products, average dot product, forward FFT, backward FFT,
 * contraction and apply, numbers of matrix-vector products and dot
 * products; Integral_Equation_Module.F90:473-484, ie_solver_module.F90:
 * 98-113) and resets them.  Times are HIP-event times of the applies and
 * Gram-Schmidt passes made while profiling was on (giem2g_op_set_profiling)
 * or timed by giem2g_op_time_apply. */
void giem2g_print_stats(void *ie_op);

/* Replaces GIEM2G_DELETE_OPERATOR (giem2g_interfaces.F90:248-254): frees
 * the device buffers and the object. */
void giem2g_delete_ie_operator(void *ie_op);

/* ===================================================================== */
/* 2. Extensions (no reference counterpart).  Return 0 on success, a     */
/*    negative code on error; giem2g_last_error() gives the message.     */
/* ===================================================================== */
#define GIEM2G_ORDER_XYZ 0
#define GIEM2G_ORDER_YXZ 1

const char *giem2g_last_error(void);

/* Component slot order of in/out vectors (default GIEM2G_ORDER_XYZ). */
int giem2g_set_component_order(void *ie_op, int order);

/* Status-returning creation without the reference structs. */
int giem2g_op_create(int Nx, int Ny, int Nz, double dx, double dy,
                     const double *z /* Nz+1 */, void **ie_op);
int giem2g_op_destroy(void *ie_op); /* fails while receiver operators use it */
/* csigb: 2*(Nl+1) doubles (complex, [0] = air); thick: Nl-1 doubles. */
int giem2g_op_set_background(void *ie_op, int Nl, const double *csigb,
                             const double *thick, double omega);
/* Tensor build + tensor spectrum (giem2g_calc_ie_kernel followed by
 * giem2g_calc_fft_of_ie_kernel). */
int giem2g_op_build_tensor(void *ie_op);
/* Only the real-space tensor build (kind 0) or only its spectrum (kind 1). */
int giem2g_op_tensor_stage(void *ie_op, int kind);
/* Copy the current tensor (real-space or spectral, Nx*Ny blocks of
 * 4*Nz*(Nz+1)/2 + 2*Nz*Nz complex, block b = kx + Nx*ky) to/from the host. */
int giem2g_op_get_tensor(void *ie_op, double *G_host);
int giem2g_op_put_tensor(void *ie_op, const double *G_host, int spectral);
/* Copy the complex (Nx,Ny,Nz) conductivity from host memory / HBM.  The
 * per-cell scalings are derived from it and the background; a later
 * background change derives them again from this copy. */
int giem2g_op_set_sigma(void *ie_op, const double *csiga_host);
int giem2g_op_set_sigma_dev(void *ie_op, const void *csiga_dev);
/* New cell sizes dx, dy and layer borders z (Nz+1) for the same Nx, Ny, Nz
 * (what giem2g_calc_ie_kernel takes from its anomaly argument): Lnumber and
 * the scalings follow; the tensor must be built again. */
int giem2g_op_set_geometry(void *ie_op, double dx, double dy, const double *z);
/* out = A*in on HBM pointers, complex128 (the FP64 drop-in) or complex64. */
int giem2g_op_apply_dev(void *ie_op, const void *in_dev, void *out_dev);
int giem2g_op_apply_dev_f32(void *ie_op, const void *in_dev, void *out_dev);
/* out0 = A*in0 and out1 = A*in1 with one read of the tensor (two-column
 * matrix-core contraction; needs Nz from 8 to 64); precision 0 complex128,
 * 1 complex64.  Under mode 0 with the x-y mirror at Nz where mode 5 serves
 * one vector, two complex128 vectors take one mirrored contraction each
 * (each reads half the tensor; together faster than the two-column kernel).  No
 * reference counterpart: the reference applies the operator to one vector
 * at a time (apply_ie_operator_module.F90:39-96). */
int giem2g_op_apply_pair_dev(void *ie_op, int precision, const void *in0_dev, const void *in1_dev,
                             void *out0_dev, void *out1_dev);
/* out[r] = A*in[r] for nrhs vectors stored one after another (3*Nx*Ny*Nz
 * complex each, complex128 for precision 0, complex64 for 1).  One tensor
 * read serves up to 8 complex64 or 4 complex128 vectors at Nz 8 to 32 and
 * 2 at Nz 33 to 64 (the matrix-core contraction with one 16-column operand
 * per two vectors; a layer count that is not a multiple of 16 is padded with
 * zero layers inside the kernel); Nz below 8 applies the vectors one by one.  No reference counterpart (the
 * reference applies the operator to one vector at a time). */
int giem2g_op_apply_batch_dev(void *ie_op, int precision, int nrhs, const void *in_dev, void *out_dev);
/* Allocation-free variant of giem2g_op_apply_batch_dev, with the same
 * arguments and products.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The call allocates and
 * frees nothing.  Its FFT intermediate and spectrum slots live in a file-scope
 * device array of 256 MiB per GPU, held from library load to process end, and
 * as many vectors share one tensor read as fit in it (128x128x32: two
 * complex128 or four complex64).  A grid whose single vector does not fit
 * (for example 256x256x32 complex128) is rejected with -3, and so is a call
 * with the profiling mode on.  For precision 1 the complex64 operator must
 * already exist (any earlier complex64 apply makes it); otherwise -4.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (two operators on different streams) race. */
int giem2g_op_apply_batch_dev_noalloc(void *ie_op, int precision, int nrhs, const void *in_dev,
                                      void *out_dev);
/* Same on host pointers (PCIe copies included). */
int giem2g_op_apply_host(void *ie_op, const double *in, double *out);
/* Stream used by the operator (hipStream_t) and device synchronisation. */
int giem2g_op_stream(void *ie_op, void **stream);
int giem2g_op_sync(void *ie_op);
/* Counters: [matvecs, matvecs32, dotprods, t_apply, t_fwd, t_contr, t_bwd,
 * t_dot, t_tensor, t_tensor_fft, device_bytes] */
int giem2g_op_stats(void *ie_op, double *out11);
/* The first n (at most 20) of: the 11 values of giem2g_op_stats, then
 * Gram-Schmidt passes, bytes moved by the Krylov kernels, FP64 contraction
 * launches, right-hand sides they served, their time (s), the same three for
 * complex64, and 1 while profiling is on.  Times in seconds. */
int giem2g_op_counters(void *ie_op, double *out, int n);
int giem2g_op_reset_counters(void *ie_op);
/* on != 0: record HIP events around the stages of every apply (forward
 * FFT, contraction, inverse FFT) and around every Gram-Schmidt pass, on the
 * streams they run on; the times reach the counters at the next query.
 * Costs a few microseconds of host time per apply. */
int giem2g_op_set_profiling(void *ie_op, int on);
/* Per-kernel timing of one apply: [fwd FFT, contraction, bwd FFT] in ms,
 * measured with HIP events on the operator stream (precision 0 FP64, 1 FP32). */
int giem2g_op_time_apply(void *ie_op, int precision, const void *in_dev,
                         void *out_dev, int reps, double *ms3);

/* Plane-wave background field (PlaneWaveIntegral, sources_module.F90:67-101)
 * for polarization pol (0 = EX, 1 = EY) into a complex (Nx,Ny,Nz,3) host
 * array, using the operator's current background and geometry. */
int giem2g_plane_wave(void *ie_op, int pol, double *E_host);

/* Nested FGMRES solve (SolveEquation, ie_solver_module.F90:43-114 with
 * GFGMRES/gmres.F90): outer flexible FGMRES(outer_m) in FP64, inner
 * GMRES(inner_m) as the right preconditioner, restarts up to max_restart,
 * tolerance on the relative residual of the scaled system.
 *   inner_precision: 0 = FP64 inner (reference semantics), 1 = complex64
 *                    inner operator and Krylov basis (FP32 accumulation)
 *   precond:         0 = inner GMRES (reference semantics);
 *                    1 = learned start: c = M(v) scaled by the least-
 *                        squares alpha, accepted when ||v - alpha A c|| <
 *                        accept*||v|| (inner precision), then z = alpha c +
 *                        inner GMRES on v - alpha A c; else plain inner GMRES;
 *                    2 = learned step only: z = M(v), accepted when the
 *                        outer's own A z gives min_alpha ||v - alpha A z|| <
 *                        accept*||v||, else z = inner GMRES(v).
 *                    accept < 1 (0.9 by default in the Python host): the
 *                    least-squares residual never exceeds ||v||.
 *                    Modes 1 and 2 need giem2g_op_set_learned. */
typedef struct {
  int outer_m;
  int inner_m;
  int max_restart;
  double tol;
  int inner_precision;
  int precond;
  int verbose;
  int learned_patience; /* learned modes: after this many rejected learned
                           steps in a row the network is not called again in
                           the solve; 0 = 2 (default), < 0 = never stop */
  double inner_tol;     /* inner solves stop when their residual falls below
                           inner_tol times their right-hand side's norm;
                           < 0 = relaxed: |inner_tol| * tol / (the outer's
                           current relative residual), between tol and 0.1
                           (inexact Krylov, Simoncini and Szyld 2003);
                           0 = default: tol for FP64 inner solves (the
                           reference's inner and outer share one tolerance,
                           seq_fgmres.F90:48-49, 62-65), -1 for complex64
                           inner solves (inner_precision 1) */
} giem2g_solver_params;

typedef struct {
  int outer_iterations; /* Arnoldi steps of the outer solver */
  int status;           /* 1 converged, 2 Arnoldi-converged only, -1 not */
  double arnoldi_residual;
  double true_residual; /* FP64 ||b - A x|| / ||b|| */
  long long matvecs_fp64;
  long long matvecs_fp32;
  long long precond_calls;
  long long precond_fallbacks;
  double seconds;
} giem2g_solve_info;

/* E_bkg, E0 (nullable warm start), E_sol: complex (Nx,Ny,Nz,3).  _dev takes
 * HBM pointers, the plain form host pointers.  The calling thread polls for
 * each Gram-Schmidt pass's results in pinned host memory (busy, for at most
 * GIEM2G_POLL_MS milliseconds, default 200, before a blocking wait;
 * GIEM2G_POLL_MS=0 blocks at once). */
int giem2g_solve(void *ie_op, const double *E_bkg, const double *E0,
                 const giem2g_solver_params *p, double *E_sol,
                 giem2g_solve_info *info);
int giem2g_solve_dev(void *ie_op, const void *E_bkg_dev, const void *E0_dev,
                     const giem2g_solver_params *p, void *E_sol_dev,
                     giem2g_solve_info *info);
/* Kernel of the per-(kx,ky) tensor contraction of one vector: 0 automatic
 * (with the x-y mirror 5 or 6, as below; without it 4 for FP64 at Nz from 8
 * to 32 and complex64 at Nz from 24 to 32, else 2
 * for Nz from 8 to 64, otherwise 1; complex64 products at Nz = 17-23,
 * 33-35 and 49-63 use 1, measured faster there than the zero-padded
 * matrix-core tiles), 1 vector ALU with two tensor components in flight,
 * 2 matrix cores (MFMA; Nz from 8 to 64, padded to a multiple of 16), 3
 * vector ALU with one component in flight, 4 small-block matrix cores (Nz
 * from 8 to 32, padded to 16 or 32; the four points are the operand columns
 * of 4x4 MFMA blocks), 5 the same with the x-y mirror: one tensor read for
 * the quads (kx, ky) and (ky, kx), exact for a layered earth on a square
 * grid with dx = dy (needs Nx = Ny, dx = dy and a tensor built by the
 * library; automatic under mode 0 where these hold; giem2g_op_put_tensor or
 * a new geometry turns it off until the next build; batched applies read
 * the full tensor).  All compute the same product, the mirror within the
 * tensor's own x-y asymmetry (about 1e-13 relative); 6 the 16x16x4 tiles
 * with the x-y mirror (Nz from 8 to 64; the mirrored quad fills the operand
 * columns a one-vector tile would repeat).  Under mode 0 with the mirror:
 * mode 5 for FP64 at Nz 8-32 and complex64 at Nz 24-32, mode 6 otherwise.
 * Two-vector products (giem2g_op_apply_pair_dev, giem2g_solve_pair_dev) use
 * the matrix cores: in complex128 two mode-5 contractions under mode 0 where
 * mode 5 would serve one vector, else the two-column tiles. */
int giem2g_op_set_contraction(void *ie_op, int mode);

/* Scalings between the field and the scaled system (ie_solver_module.F90:
 * 64-66, 85-97) on complex128 HBM vectors: mode 0 out = sqrt(Re sb)*in (the
 * right-hand side from E_bkg), 1 out = in/a (a field to the unknown x),
 * 2 out = a*in (x to the field), a = 2 sqrt(Re sb)/(sa + conj sb). */
int giem2g_op_scale(void *ie_op, int mode, const void *in_dev, void *out_dev);

/* Two independent solves with the same operator (e.g. the EY and EX
 * polarizations of one frequency) run together: their operator products
 * are paired so that one two-vector apply reads the tensor once for both
 * (needs Nz from 8 to 64, otherwise the products run one by one).  HBM
 * pointers; E0x may be NULL; info2 receives two records (seconds = wall
 * time of the pair); paired (nullable) receives [two-vector applies,
 * one-vector applies].  precond must be 0.  No reference counterpart: the
 * reference solves the polarizations one after the other
 * (giem2g.F90:346, 435). */
int giem2g_solve_pair_dev(void *ie_op, const void *E_bkg0_dev, const void *E_bkg1_dev,
                          const void *E00_dev, const void *E01_dev, const giem2g_solver_params *p,
                          void *E_sol0_dev, void *E_sol1_dev, giem2g_solve_info *info2,
                          long long *paired);

/* Residual history of the last solve (outer Arnoldi residuals). */
int giem2g_last_history(void *ie_op, double *out, int maxlen, int *len);

/* ---- learned right preconditioner (no reference counterpart: the
 * reference preconditions only with the inner GMRES of
 * GFGMRES/seq_fgmres.F90:142-152) ----------------------------------------
 * A 3-level 3D U-Net (widths 16/32/64, 3x3x3 convolutions, BF16 weights and
 * activations on MFMA, FP32 accumulation).  Volumes are channels-last FP32
 * arrays [B][Nz][Ny][Nx][C] in HBM: input C = 8 (residual re/im of x, y, z
 * divided by its RMS, log10(Re sa / Re sb), log10(|k_b| dx) with
 * |k_b|^2 = omega mu0 |sb|), output C = 6 (the correction a*e, re/im of
 * x, y, z, in the residual's units).  Nx, Ny, Nz must be multiples of 4. */
int giem2g_unet_create(void **net);
int giem2g_unet_destroy(void *net);
int giem2g_unet_nparams(long long *n);
/* FP32 parameters, layer by layer W[cout][27 or 1][cin] then b[cout];
 * setting them resets the Adam state. */
int giem2g_unet_set_params(void *net, const float *host);
int giem2g_unet_get_params(void *net, float *host);
int giem2g_unet_get_grads(void *net, float *host);
/* y = net(x); keep_for_backward keeps the activations for one backward. */
int giem2g_unet_forward_dev(void *net, const float *x_dev, float *y_dev, int B, int nx,
                            int ny, int nz, int keep_for_backward);
/* Parameter gradient of sum(dy * y) for the last kept forward. */
int giem2g_unet_backward_dev(void *net, const float *dy_dev);
/* Forward, MSE loss against t [B][vox][6], backward and one Adam step. */
int giem2g_unet_train_step(void *net, const float *x_dev, const float *t_dev, int B, int nx,
                           int ny, int nz, float lr, double *loss);
/* One Adam step (beta 0.9, 0.999) on the gradient of the last backward. */
int giem2g_unet_adam(void *net, float lr);
/* out = A^H in on complex128 HBM vectors (no reference counterpart; used to
 * train the network on the residual ||r - A M(r)||).  From the reciprocity
 * of the contraction kernel: diag(dz) K is complex-symmetric. */
int giem2g_op_apply_adjoint_dev(void *ie_op, const void *in_dev, void *out_dev);
/* Attach a network to an operator (net NULL detaches); accept is the
 * residual-reduction threshold of precond modes 1 and 2. */
int giem2g_op_set_learned(void *ie_op, void *net, double accept);
/* Network input x [vox][8] of a residual r and, when e is not NULL, the
 * training target t [vox][6] of the correction e (complex128 HBM vectors of
 * the operator's layout). */
int giem2g_op_learned_features(void *ie_op, const void *r_dev, const void *e_dev, float *x_dev,
                               float *t_dev);
/* z = M(v) on complex128 HBM vectors. */
int giem2g_op_learned_apply(void *ie_op, const void *v_dev, void *z_dev);

/* ---- continuation of the solution to receivers (the post-solve field
 * recalculation of the reference CLI, Continuation_Function_Module.F90,
 * calc_rc_tensor_module.F90, rc_kernel_hankel_module.F90).  A receiver
 * operator shares the grid, background (frequency) and stream of an IE
 * operator, which must outlive it.  Receivers sit at every cell centre of
 * the Nx x Ny grid shifted by (xshift, yshift) metres, at the given depth;
 * Ea, Ha are complex (Nr,3,Nx,Ny) in Fortran order with components in
 * (x,y,z) order, the anomalous field (SaveOutputOneFile adds the normal
 * field of giem2g_rc_plane_wave). ---- */
/* Replaces PrepareContinuationOperator (Continuation_Function_Module.F90:
 * 82-167) with LoadRecievers' receivers (mpi_saveload_module.F90:174-213);
 * 1 <= Nr <= 64.  Receivers on a cell corner are rejected (the reference's
 * filter weights take log(0) there). */
int giem2g_rc_create(void *ie_op, int Nr, const double *xshift, const double *yshift,
                     const double *depth, void **rc_op);
/* Replaces DeleteRC_OP (Continuation_Function_Module.F90:396-405). */
int giem2g_rc_destroy(void *rc_op);
/* Replaces CalcRecalculationGreenTensor (calc_rc_tensor_module.F90:37-162)
 * for the IE operator's current background: the real-space tensor. */
int giem2g_rc_calc_kernel(void *rc_op);
/* Replaces CalcFFTofRCTensor (Continuation_Function_Module.F90:349-383). */
int giem2g_rc_calc_fft_of_kernel(void *rc_op);
/* Copy the current receiver tensor (real space after calc_kernel, spectrum
 * after calc_fft_of_kernel): 4*Nx*Ny points p = ix*2Ny + iy, each 15*Nr*Nz
 * complex with entry e (REXX, REXY, REXZ, REYY, REYZ, REZX, REZY, REZZ,
 * RHXX, RHXY, RHXZ, RHYX, RHYZ, RHZX, RHZY), receiver r, source layer z at
 * (e*Nr + r)*Nz + z. */
int giem2g_rc_get_tensor(void *rc_op, double *G_host);
/* Load a tensor of that layout (spectral != 0: a spectrum) for the current
 * frequency. */
int giem2g_rc_put_tensor(void *rc_op, const double *G_host, int spectral);
/* PrepareRecvs (data_types_module.F90:124-152) for the current background:
 * depth after the move to a cell border, background layer, anomaly cell
 * index (0 above, Nz+1 below).  Any output may be NULL. */
int giem2g_rc_receivers(void *rc_op, double *zr, int *layer, int *cell);
/* Replaces FinalizeRCOperator (Continuation_Function_Module.F90:168-196):
 * complex (Nx,Ny,Nz) anomaly conductivity of the current frequency
 * (including -i w eps0, like giem2g_set_anomaly_conductivity). */
int giem2g_rc_set_anomaly_conductivity(void *rc_op, const double *csiga_host);
int giem2g_rc_set_anomaly_conductivity_dev(void *rc_op, const void *csiga_dev);
/* Replaces ReCalculation (Continuation_Function_Module.F90:198-240): the
 * anomalous E and H at the receivers from the total field E_int (complex
 * (Nx,Ny,Nz,3) in the IE operator's component order) inside the anomaly.
 * _dev: HBM pointers, ordered on the IE operator's stream. */
int giem2g_rc_recalculate(void *rc_op, const double *E_int, double *Ea, double *Ha);
int giem2g_rc_recalculate_dev(void *rc_op, const void *E_int_dev, void *Ea_dev, void *Ha_dev);
/* Replaces PlaneWave (sources_module.F90:26-66) at the prepared receiver
 * depths for pol 0 (EX) or 1 (EY): complex (Nr,6) Ex,Ey,Ez,Hx,Hy,Hz. */
int giem2g_rc_plane_wave(void *rc_op, int pol, double *F0_host);
/* [tensor build s, tensor FFT s, last recalculation ms, of which forward
 * FFT ms, contraction ms, inverse FFT ms, recalculations, HBM bytes]. */
int giem2g_rc_stats(void *rc_op, double *out8);

#ifdef __cplusplus
}
#endif
#endif
