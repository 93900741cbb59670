/*
 * giem2g_oracle.h -- CPU FP64 oracle for the GIEM2G integral-equation solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (DarthLaran/GIEM2G, Fortran) used as the checker for the HIP
 * implementation.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product library (giem2g_amd/) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Layout
 * conventions (shared with the device library so results can be compared
 * element by element):
 *   field vector v  : complex128, Fortran order (Nx,Ny,Nz,3):
 *                     v[x + Nx*(y + Ny*(z + Nz*c))], c = 0,1,2 = x,y,z
 *   csiga           : complex128 (Nx,Ny,Nz): x + Nx*(y + Ny*z)
 *   tensor block    : one (lx,ly) [real space] or (kx,ky) [spectrum] block of
 *                     BS = 4*P + 2*Nz*Nz complex, P = Nz*(Nz+1)/2:
 *                       [xx P][xy P][yy P][zz P]  packed upper, (i<=j) at i + j*(j+1)/2
 *                       [xz Nz*Nz][yz Nz*Nz]      column-major, (i,j) at i + Nz*j
 *                     (i = receiver layer, j = source layer).  Blocks are
 *                     ordered b = lx + Nx*ly (x fastest).
 */
#ifndef GIEM2G_ORACLE_H
#define GIEM2G_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXL 64
#define OR_NFIRST 212
#define OR_NLAST 712
#define OR_NTAPS 501

typedef struct or_model or_model;

/* Background + anomaly geometry for one frequency.
 *   csigma : 2*(Nl+1) doubles (re,im) -- [0] is air, already including -i*w*eps0
 *   thick  : Nl-1 layer thicknesses (m)
 *   omega  : angular frequency
 *   z      : Nz+1 layer borders of the anomaly (m, positive down)          */
or_model *or_model_new(int Nl, const double *csigma, const double *thick,
                       double omega, int Nx, int Ny, int Nz, double dx,
                       double dy, const double *z);
void or_model_free(or_model *m);
int or_model_lnum(const or_model *m, int iz); /* iz in 0..Nz (0 = top border) */

/* Physical constants of the reference (const_module.F90:31-35). */
double or_mu0(void);
double or_eps0(void);

/* Hankel digital filter (VolumeBesselTransforms.F90). */
void or_vbt_lms(double rmin, double *lms /*501*/);
void or_vbt_weights(double x, double y, double hx, double hy, double rmin,
                    double *WT /*501*6, (K,i) column-major: K + 501*i*/);
void or_compute_weights(int lx, int ly, double dx, double dy, double rmin,
                        double *WT /*501*6 in IE order DXX,DXY,DYY,DX,DY,D0*/,
                        double *W0 /*6*/);

/* Layered kernel for one lambda (apq_module.F90, ie_kernel_hankel_module.F90). */
void or_calc_apq(const or_model *m, double lm, double *Arr, double *p,
                 double *q, double *eta, double *A /*nullable*/);
void or_double_integral_U(const or_model *m, double lm, const double *W /*6*/,
                          double *G /*complex 6*Nz*Nz: c + 6*(i + Nz*j)*/);

/* Real-space tensor (calc_ie_tensor_module.F90). */
int64_t or_block_size(int Nz);
void or_tensor_element(const or_model *m, int lx, int ly, double *blk);
/* per entry, sum over taps K and weights c of max|WT_c| |dG/dWT_c(K)|: the change
   of or_tensor_element per weight error of one unit of each vector's largest weight */
void or_tensor_element_abssum(const or_model *m, int lx, int ly, double *out);
void or_tensor_realspace(const or_model *m, double *Greal, int nthreads);

/* Tensor spectrum (Integral_Equation_Module.F90:319-471). */
void or_tensor_fft(const or_model *m, const double *Greal, double *Gspec,
                   int nthreads);

/* Operator apply (apply_ie_operator_module.F90:39-374). */
void or_apply(const or_model *m, const double *Gspec, const double *csiga,
              const double *vin, double *vout, int nthreads);
void or_apply_bruteforce(const or_model *m, const double *Greal,
                         const double *csiga, const double *vin, double *vout,
                         int nthreads);
/* Split timings of the last or_apply call: fwd FFT, contraction, bwd FFT (s). */
void or_apply_timings(double *t3);

/* Plane-wave RHS (sources_module.F90:67-101); pol 0 = EX, 1 = EY. */
void or_plane_wave_integral(const or_model *m, int pol, double *E0);

/* Nested FGMRES (GFGMRES/gmres.F90, GFGMRES/seq_fgmres.F90). */
typedef void (*or_matvec_fn)(void *ctx, const double *in, double *out);
typedef struct {
  int iterations; /* outer Arnoldi steps */
  int stat;       /* 1 converged, 2 non-true-converged, -1 non-converged */
  double bea;     /* Arnoldi residual */
  double be;      /* true residual */
  int64_t matvecs;
  int64_t dotprods;
} or_gmres_info;
void or_fgmres(int64_t N, int m_out, int m_in, int maxit, double tol,
               or_matvec_fn A, void *ctx, const double *x0, const double *b,
               double *x, or_gmres_info *info, double *bea_hist,
               int hist_len);

/* SolveEquation (ie_solver_module.F90:43-114) on the FFT operator. */
void or_solve(const or_model *m, const double *Gspec, const double *csiga,
              const double *Ebkg, const double *E0 /*nullable*/, int m_out,
              int m_in, int maxit, double tol, double *Esol,
              or_gmres_info *info, int nthreads);

/* Receiver recalculation (Continuation_Function_Module.F90,
 * calc_rc_tensor_module.F90, rc_kernel_hankel_module.F90).
 *   receivers     : xs, ys lateral shifts (m) from the cell centres, zr depths
 *                   after or_prepare_recvs, layer = background layer,
 *                   cell = anomaly cell index (0 above, Nz+1 below)
 *   RC tensor     : per offset point p = ix*2Ny + iy of the 2Nx x 2Ny grid a
 *                   block of 15*Nr*Nz complex; entry e (0..7 = REXX, REXY,
 *                   REXZ, REYY, REYZ, REZX, REZY, REZZ; 8..14 = RHXX, RHXY,
 *                   RHXZ, RHYX, RHYZ, RHZX, RHZY), receiver r, source layer z
 *                   at (e*Nr + r)*Nz + z
 *   Ea, Ha        : complex (Nr,3,Nx,Ny) Fortran order                       */
void or_prepare_recvs(const or_model *m, int Nr, const double *depth,
                      double *zr, int *layer, int *cell);
void or_vbt_weights_single(double x, double y, double hx, double hy,
                           double rmin, double *WT /*501*6 (K,i), D0,DXX,DXY,DX,DYY,DY*/);
void or_rc_integral_U(const or_model *m, double zrecv, int layer, int cell,
                      double lm, const double *W /*6 IE order*/,
                      double *GE /*complex 8*Nz*/, double *GH /*complex 7*Nz*/);
int64_t or_rc_block_size(int Nz, int Nr);
void or_rc_tensor_element(const or_model *m, int Nr, const double *xs,
                          const double *ys, const double *zr, const int *layer,
                          const int *cell, int lx, int ly, double *blk);
void or_rc_tensor_realspace(const or_model *m, int Nr, const double *xs,
                            const double *ys, const double *zr,
                            const int *layer, const int *cell, double *Greal,
                            int nthreads);
void or_rc_tensor_fft(const or_model *m, int Nr, const double *Greal,
                      double *Gspec, int nthreads);
void or_rc_apply(const or_model *m, int Nr, const double *Gspec,
                 const double *csiga, const double *Eint, double *Ea,
                 double *Ha, int nthreads);
void or_rc_apply_bruteforce(const or_model *m, int Nr, const double *Greal,
                            const double *csiga, const double *Eint,
                            double *Ea, double *Ha, int nthreads);
/* PlaneWave (sources_module.F90:26-66); F0 complex (Nr,6) Ex,Ey,Ez,Hx,Hy,Hz */
void or_plane_wave(const or_model *m, int pol, int Nr, const double *zrecv,
                   double *F0);

/* Radix-2 complex FFT, length n (power of two), in place; sign -1 forward. */
void or_fft(double *a, int n, int sign);

#ifdef __cplusplus
}
#endif
#endif
