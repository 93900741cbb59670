/*
 * giem2g_oracle.c -- CPU FP64 restatement of the GIEM2G integral-equation
 * solve (reference: DarthLaran/GIEM2G, Fortran 2003).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP library.
 * It is loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU
 * baseline leg, never by the product.  See giem2g_oracle.h for layouts.
 *
 * Pins (what makes each part trustworthy beyond "retyped formula"):
 *  - Hankel weights / layered kernel / FGMRES are compared against the
 *    reference's own Fortran sources compiled into oracle/_ref (see
 *    oracle/ref_harness.F90 and oracle/Makefile), when that builds.
 *  - The FFT operator is compared against or_apply_bruteforce (direct
 *    block-Toeplitz sum, no FFT, no Nyquist tricks) and against the
 *    reciprocity / identity properties in tests/test_oracle.py.
 */
#include "giem2g_oracle.h"

#include <complex.h>
#undef I
#define CI _Complex_I
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* const_module.F90:31-35 */
static const double PI = 3.1415926535897932384626433832795028841971693993751;
#define MU0 (4.0 * PI * 1e-7)
#define EPS0 (1.0 / (MU0 * 299792458.0 * 299792458.0))

double or_mu0(void) { return MU0; }
double or_eps0(void) { return EPS0; }

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Model: BKG_DATA_TYPE + ANOMALY_TYPE (data_types_module.F90:23-41)        */
/* ------------------------------------------------------------------------ */
struct or_model {
  int Nl;
  double omega;
  cplx iwm;
  cplx csigma[OR_MAXL + 1]; /* 0..Nl, 0 = air */
  cplx k2[OR_MAXL + 1];
  double thick[OR_MAXL + 1]; /* 1..Nl-1 */
  double depth[OR_MAXL + 1]; /* 1..Nl-1 */
  int Nx, Ny, Nz;
  double dx, dy;
  double *z;  /* 0..Nz */
  double *dz; /* 1..Nz */
  int *lnum;  /* 0..Nz */
};

/* GetLayerSingle (data_types_module.F90:75-103) */
static int get_layer(const or_model *m, double z) {
  int l;
  if (fabs(z) < 1e-15) return 0;
  if (z < 0) return -1;
  if (m->Nl == 1) return 1;
  l = 1;
  if (z > m->depth[m->Nl - 1]) return m->Nl;
  while (z > m->depth[l]) l++;
  return l;
}

or_model *or_model_new(int Nl, const double *csigma, const double *thick,
                       double omega, int Nx, int Ny, int Nz, double dx,
                       double dy, const double *z) {
  int i;
  or_model *m;
  if (Nl < 1 || Nl > OR_MAXL || Nz < 1) return NULL;
  m = (or_model *)calloc(1, sizeof(or_model));
  m->Nl = Nl;
  m->omega = omega;
  m->iwm = CI * MU0 * omega; /* data_types_module.F90:119 */
  for (i = 0; i <= Nl; i++) {
    m->csigma[i] = csigma[2 * i] + CI * csigma[2 * i + 1];
    m->k2[i] = CI * m->csigma[i] * MU0 * omega; /* :117 */
  }
  /* LoadBackground (mpi_saveload_module.F90:55-61) */
  for (i = 1; i <= Nl - 1; i++) m->thick[i] = thick[i - 1];
  if (Nl > 1) {
    m->depth[1] = m->thick[1];
    for (i = 2; i <= Nl - 1; i++) m->depth[i] = m->depth[i - 1] + m->thick[i];
  }
  m->Nx = Nx;
  m->Ny = Ny;
  m->Nz = Nz;
  m->dx = dx;
  m->dy = dy;
  m->z = (double *)malloc(sizeof(double) * (Nz + 1));
  m->dz = (double *)malloc(sizeof(double) * (Nz + 1));
  m->lnum = (int *)malloc(sizeof(int) * (Nz + 1));
  for (i = 0; i <= Nz; i++) m->z[i] = z[i];
  /* LoadAnomalyShape (mpi_saveload_module.F90:101-103) */
  for (i = 1; i <= Nz; i++) {
    m->lnum[i] = get_layer(m, (m->z[i - 1] + m->z[i]) / 2.0);
    m->dz[i] = m->z[i] - m->z[i - 1];
  }
  m->lnum[0] = get_layer(m, m->z[0] * (1.0 - 1e-7));
  m->dz[0] = 0;
  return m;
}

void or_model_free(or_model *m) {
  if (!m) return;
  free(m->z);
  free(m->dz);
  free(m->lnum);
  free(m);
}

int or_model_lnum(const or_model *m, int iz) { return m->lnum[iz]; }

/* ------------------------------------------------------------------------ */
/* Radix-2 FFT (own implementation; any exact DFT reproduces the reference's */
/* FFTW / Ooura transforms to rounding).                                     */
/* ------------------------------------------------------------------------ */
static void dft_naive(cplx *a, int n, int sign) {
  cplx *t = malloc(sizeof(cplx) * n);
  int k, j;
  for (k = 0; k < n; k++) {
    cplx acc = 0;
    for (j = 0; j < n; j++) {
      double ang = sign * 2.0 * PI * (double)(((int64_t)j * k) % n) / (double)n;
      acc += a[j] * (cos(ang) + CI * sin(ang));
    }
    t[k] = acc;
  }
  memcpy(a, t, sizeof(cplx) * n);
  free(t);
}

static void fft_c(cplx *a, int n, int sign) {
  int i, j, k, len;
  if (n & (n - 1)) { /* not a power of two: plain DFT */
    dft_naive(a, n, sign);
    return;
  }
  for (i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      cplx t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  for (len = 2; len <= n; len <<= 1) {
    int half = len >> 1;
    for (k = 0; k < half; k++) {
      double ang = sign * 2.0 * PI * (double)k / (double)len;
      cplx w = cos(ang) + CI * sin(ang);
      for (i = k; i < n; i += len) {
        cplx u = a[i], v = a[i + half] * w;
        a[i] = u + v;
        a[i + half] = u - v;
      }
    }
  }
}

void or_fft(double *a, int n, int sign) { fft_c((cplx *)a, n, sign); }

/* ------------------------------------------------------------------------ */
/* Hankel digital filter: VolumeBesselTransforms.F90 (FFT_QUAD_OOURA branch) */
/* ------------------------------------------------------------------------ */
#define VB_N 1024
#define VB_N2 (-512)
static const double VB_YSTEP = 1e-1 * 2e0;                      /* :41 */
static const double VB_XI = -4.82205425726955430357775925664462079e-2 * 2e0; /* :42 */
#define VB_Q (VB_YSTEP / 2.0)                                    /* :43 */
#define VB_P (VB_Q + VB_XI)                                      /* :44 */
/* Ooura cdft convention: FWD = +1 (exp(+i)), BWD = -1 (:56-57). */
#define VB_FWD (+1)
#define VB_BWD (-1)

/* inputfunction (:292-300) */
static double vb_inputfunction(double t) {
  double y = exp(-t), y2 = y * y, y4 = y2 * y2;
  return 8.0 * exp(-y2) * (y4 - 4.0 * y2 + 2.0) * y;
}

/* VBTransformInit (:89-109): lms and the spectrum f1 of the input function */
static void vb_init(double s, double *lms, cplx *f1) {
  int i, j;
  for (j = OR_NFIRST; j <= OR_NLAST; j++)
    lms[j - OR_NFIRST] = exp(VB_XI + (double)(j + VB_N2) * VB_YSTEP) / s;
  for (i = 0; i < VB_N; i += 2) {
    double y1 = (double)(i + VB_N2) * VB_YSTEP + VB_Q;
    f1[i] = vb_inputfunction(y1);
    y1 = y1 + VB_YSTEP;
    f1[i + 1] = -vb_inputfunction(y1);
  }
  fft_c(f1, VB_N, VB_FWD);
}

void or_vbt_lms(double rmin, double *lms) {
  int j;
  for (j = OR_NFIRST; j <= OR_NLAST; j++)
    lms[j - OR_NFIRST] = exp(VB_XI + (double)(j + VB_N2) * VB_YSTEP) / rmin;
}

typedef struct {
  double explt[VB_N], explt2[VB_N], explt3[VB_N];
  double xy[2][VB_N][3], xy2[2][VB_N][3], xexp2[2][VB_N][3], xerf[2][VB_N][3];
} vb_edt;

/* CALC_EXP_LT (:142-157) */
static void vb_calc_exp_lt(double s, vb_edt *e) {
  int i;
  for (i = 0; i < VB_N; i += 2) {
    double y2 = (double)(i + VB_N2) * VB_YSTEP + VB_P + s;
    e->explt[i] = exp(y2);
    e->explt2[i] = e->explt[i] * e->explt[i];
    e->explt3[i] = e->explt2[i] * e->explt[i];
    y2 = y2 + VB_YSTEP;
    e->explt[i + 1] = exp(y2);
    e->explt2[i + 1] = e->explt[i + 1] * e->explt[i + 1];
    e->explt3[i + 1] = e->explt2[i + 1] * e->explt[i + 1];
  }
}

/* PrepareExps4 (:304-340) */
static void vb_prepare_exps4(double phi, double alpha, double beta, vb_edt *e) {
  int i, k;
  double cphi = cos(phi), sphi = sin(phi);
  for (i = 0; i < VB_N; i++) {
    double x1 = e->explt[i] * cphi, y1 = e->explt[i] * sphi;
    double ddx = alpha * e->explt[i], ddy = beta * e->explt[i];
    double x[3], y[3];
    x1 = x1 - ddx * 0.5;
    y1 = y1 - ddy * 0.5;
    x[0] = x1;
    x[1] = x[0] + ddx;
    x[2] = x[0] - ddx;
    y[0] = y1;
    y[1] = y[0] + ddy;
    y[2] = y[0] - ddy;
    for (k = 0; k < 3; k++) {
      double xv = x[k] * 0.5, yv = y[k] * 0.5;
      e->xy[0][i][k] = xv;
      e->xy2[0][i][k] = xv * xv;
      e->xexp2[0][i][k] = exp(-(xv * xv));
      e->xerf[0][i][k] = erf(xv);
      e->xy[1][i][k] = yv;
      e->xy2[1][i][k] = yv * yv;
      e->xexp2[1][i][k] = exp(-(yv * yv));
      e->xerf[1][i][k] = erf(yv);
    }
  }
}

/* second difference f(2)+f(3)-2 f(1) used by every INT_* routine */
#define D2(f) ((f)[1] + (f)[2] - 2.0 * (f)[0])

/* INT_2d_edt (:533-550) and DOUBLE_INT_D0 (:509-518) */
static void vb_double_int_d0(const vb_edt *e, int i, int j, double *R) {
  const double *x = e->xy[j][i], *x2 = e->xy2[j][i];
  double f[3], r[3];
  int k;
  for (k = 0; k < 3; k++) f[k] = x[k] * e->xerf[j][i][k];
  r[0] = D2(f) * sqrt(PI) * 0.25;
  for (k = 0; k < 3; k++) f[k] = e->xexp2[j][i][k];
  r[1] = D2(f);
  for (k = 0; k < 3; k++) f[k] = e->xexp2[j][i][k] * x2[k] / 4.0;
  r[2] = D2(f);
  R[0] = 32.0 * (1.5 * r[0] + r[1] + r[2]);
  R[1] = 8.0 * r[0] + 4.0 * r[1];
  R[2] = 4.0 * r[0] + r[1];
}

/* INT_1d_edt (:493-508) and DOUBLE_INT_D1 (:470-480) */
static void vb_double_int_d1(const vb_edt *e, int i, int j, double *R) {
  const double *x = e->xy[j][i], *x2 = e->xy2[j][i];
  double f[3], r[3];
  int k;
  for (k = 0; k < 3; k++) f[k] = e->xerf[j][i][k];
  r[0] = D2(f) * sqrt(PI) * 0.5;
  for (k = 0; k < 3; k++) f[k] = e->xexp2[j][i][k] * x[k];
  r[1] = D2(f);
  for (k = 0; k < 3; k++) f[k] = f[k] * x2[k];
  r[2] = D2(f);
  R[0] = -8.0 * r[2] - 12.0 * r[1] + 12.0 * r[0];
  R[1] = -2.0 * r[1] + 2.0 * r[0];
  R[2] = r[0];
}

/* DOUBLE_INT_D2 (:437-452) */
static void vb_double_int_d2(const vb_edt *e, int i, int j, double *R) {
  const double *x2 = e->xy2[j][i];
  double ex[3], f[3];
  int k;
  for (k = 0; k < 3; k++) ex[k] = e->xexp2[j][i][k] / 2.0;
  for (k = 0; k < 3; k++) f[k] = x2[k] * ex[k] * 4.0;
  R[1] = D2(f);
  for (k = 0; k < 3; k++) f[k] = x2[k] * f[k] * 4.0;
  R[0] = D2(f);
  R[2] = D2(ex);
}

/* outfunc_int4_all (:377-403) */
static void vb_outfunc_int4(const vb_edt *e, int i, double *R) {
  double fx[3][3], fy[3][3]; /* [D0,D1,D2][1..3] */
  vb_double_int_d0(e, i, 0, fx[0]);
  vb_double_int_d1(e, i, 0, fx[1]);
  vb_double_int_d2(e, i, 0, fx[2]);
  vb_double_int_d0(e, i, 1, fy[0]);
  vb_double_int_d1(e, i, 1, fy[1]);
  vb_double_int_d2(e, i, 1, fy[2]);
#define CMB(a, b) (fx[a][2] * fy[b][0] + fx[a][0] * fy[b][2] + 2.0 * fx[a][1] * fy[b][1])
  R[0] = CMB(0, 0) / e->explt3[i];
  R[1] = CMB(2, 0) / e->explt[i];
  R[2] = CMB(1, 1) / e->explt[i];
  R[3] = CMB(1, 0) / e->explt2[i];
  R[4] = CMB(0, 2) / e->explt[i];
  R[5] = CMB(0, 1) / e->explt2[i];
#undef CMB
}

/* VBTransformWeightsAllDoubleInt (:159-197) + CalcWeights22 (:231-262).
 * WT is (K,i) column-major with K in [0,501), i in [0,6) in the reference's
 * output order (D0, DXX, DXY, DX, DYY, DY). */
void or_vbt_weights(double x, double y, double hx, double hy, double rmin,
                    double *WT) {
  double lms[OR_NTAPS];
  cplx f1[VB_N];
  cplx *g = (cplx *)malloc(sizeof(cplx) * VB_N * 6);
  vb_edt *e = (vb_edt *)malloc(sizeof(vb_edt));
  double rho, rp[6], alpha, beta, phi;
  int i, j;
  vb_init(rmin, lms, f1);
  rho = sqrt(x * x + y * y);
  rp[0] = rho * rho * rho;
  rp[1] = rho;
  rp[2] = rho;
  rp[3] = rho * rho;
  rp[4] = rho;
  rp[5] = rho * rho;
  alpha = hx / rho;
  beta = hy / rho;
  phi = atan2(y, x);
  vb_calc_exp_lt(log(rho) - log(rmin), e);
  vb_prepare_exps4(phi, alpha, beta, e);
  for (i = 0; i < VB_N; i += 2) {
    double R[6];
    vb_outfunc_int4(e, i, R);
    for (j = 0; j < 6; j++) g[j * VB_N + i] = R[j];
    vb_outfunc_int4(e, i + 1, R);
    for (j = 0; j < 6; j++) g[j * VB_N + i + 1] = -R[j];
  }
  for (j = 0; j < 6; j++) {
    cplx *h = g + j * VB_N;
    fft_c(h, VB_N, VB_FWD);
    for (i = 0; i < VB_N; i += 2) {
      h[i] = h[i] / f1[i];
      h[i + 1] = -h[i + 1] / f1[i + 1];
    }
    fft_c(h, VB_N, VB_BWD);
    for (i = 1; i < VB_N; i += 2) {
      h[i - 1] = h[i - 1] / (double)VB_N;
      h[i] = -h[i] / (double)VB_N;
    }
    for (i = OR_NFIRST; i <= OR_NLAST; i++)
      WT[(i - OR_NFIRST) + OR_NTAPS * j] = creal(h[i]) / hx / hy * rp[j];
  }
  free(e);
  free(g);
}

/* IE weight order (ie_kernel_hankel_module.F90:29-35) */
enum { IE_DXX = 0, IE_DXY = 1, IE_DYY = 2, IE_DX = 3, IE_DY = 4, IE_D0 = 5 };

static double maxabs_col(const double *WT, int col) {
  double mx = 0;
  int k;
  for (k = 0; k < OR_NTAPS; k++) {
    double a = fabs(WT[k + OR_NTAPS * col]);
    if (a > mx) mx = a;
  }
  return mx;
}

/* COMPUTE_WEIGHTS (calc_ie_tensor_module.F90:308-352) */
void or_compute_weights(int lx, int ly, double dx, double dy, double rmin,
                        double *WT, double *W0) {
  double WT0[OR_NTAPS * 6];
  double x = lx * dx + dx / 2.0, y = ly * dy + dy / 2.0;
  int k;
  or_vbt_weights(x, y, dx, dy, rmin, WT0);
  /* column map :320-325 (WT0 order D0,DXX,DXY,DX,DYY,DY) */
  memcpy(WT + OR_NTAPS * IE_D0, WT0 + OR_NTAPS * 0, sizeof(double) * OR_NTAPS);
  memcpy(WT + OR_NTAPS * IE_DXX, WT0 + OR_NTAPS * 1, sizeof(double) * OR_NTAPS);
  memcpy(WT + OR_NTAPS * IE_DXY, WT0 + OR_NTAPS * 2, sizeof(double) * OR_NTAPS);
  memcpy(WT + OR_NTAPS * IE_DX, WT0 + OR_NTAPS * 3, sizeof(double) * OR_NTAPS);
  memcpy(WT + OR_NTAPS * IE_DYY, WT0 + OR_NTAPS * 4, sizeof(double) * OR_NTAPS);
  memcpy(WT + OR_NTAPS * IE_DY, WT0 + OR_NTAPS * 5, sizeof(double) * OR_NTAPS);
  W0[IE_DXX] = maxabs_col(WT, IE_DXX);
  if (lx != 0 && ly != 0) {
    /* :329 maxval over columns DXY..DY is one scalar */
    double mx = 0;
    int c;
    for (c = IE_DXY; c <= IE_DY; c++) {
      double t = maxabs_col(WT, c);
      if (t > mx) mx = t;
    }
    for (c = IE_DXY; c <= IE_DY; c++) W0[c] = mx;
  } else if (lx == 0) {
    for (k = 0; k < OR_NTAPS; k++) WT[k + OR_NTAPS * IE_DXY] = 0;
    W0[IE_DXY] = -W0[IE_DXX];
    W0[IE_DYY] = maxabs_col(WT, IE_DYY);
    for (k = 0; k < OR_NTAPS; k++) WT[k + OR_NTAPS * IE_DX] = 0;
    W0[IE_DX] = -W0[IE_DXX];
    if (ly != 0) {
      W0[IE_DY] = maxabs_col(WT, IE_DY);
    } else {
      for (k = 0; k < OR_NTAPS; k++) WT[k + OR_NTAPS * IE_DY] = 0;
      W0[IE_DY] = -W0[IE_DXX];
    }
  } else {
    for (k = 0; k < OR_NTAPS; k++) WT[k + OR_NTAPS * IE_DXY] = 0;
    W0[IE_DXY] = -W0[IE_DXX];
    W0[IE_DYY] = maxabs_col(WT, IE_DYY);
    W0[IE_DX] = maxabs_col(WT, IE_DX);
    for (k = 0; k < OR_NTAPS; k++) WT[k + OR_NTAPS * IE_DY] = 0;
    W0[IE_DY] = -W0[IE_DXX];
  }
  W0[IE_D0] = maxabs_col(WT, IE_D0);
}

/* ------------------------------------------------------------------------ */
/* Layered kernel: apq_module.F90                                           */
/* ------------------------------------------------------------------------ */
/* Calc_p (:267-290) */
static void calc_p(const or_model *m, const cplx *eta, const cplx *el,
                   double lms, cplx (*p)[2]) {
  cplx eta0 = csqrt(lms * lms - m->k2[0]);
  cplx ceta = eta0 / eta[1], cgamma, w[2], a[2];
  int i, k;
  p[1][0] = (1.0 - ceta) / (1.0 + ceta);
  cgamma = m->csigma[0] / m->csigma[1];
  p[1][1] = (cgamma - ceta) / (cgamma + ceta);
  for (i = 2; i <= m->Nl; i++) {
    w[0] = el[i - 1] * p[i - 1][0];
    w[1] = el[i - 1] * p[i - 1][1];
    for (k = 0; k < 2; k++) w[k] = (w[k] - 1.0) / (w[k] + 1.0);
    a[0] = eta[i - 1] / eta[i];
    a[1] = a[0] * m->csigma[i] / m->csigma[i - 1];
    for (k = 0; k < 2; k++) {
      w[k] = w[k] * a[k];
      p[i][k] = (1.0 + w[k]) / (1.0 - w[k]);
    }
  }
}

/* Calc_q (:291-312) */
static void calc_q(const or_model *m, const cplx *eta, const cplx *el,
                   cplx (*q)[2]) {
  int Nl = m->Nl, i, k;
  cplx w[2], a[2];
  q[Nl][0] = q[Nl][1] = 0;
  if (Nl > 1) {
    a[0] = eta[Nl] / eta[Nl - 1];
    a[1] = a[0] * m->csigma[Nl - 1] / m->csigma[Nl];
    for (k = 0; k < 2; k++) q[Nl - 1][k] = (1.0 - a[k]) / (1.0 + a[k]);
    for (i = Nl - 2; i >= 1; i--) {
      w[0] = el[i + 1] * q[i + 1][0];
      w[1] = el[i + 1] * q[i + 1][1];
      for (k = 0; k < 2; k++) w[k] = (w[k] - 1.0) / (w[k] + 1.0);
      a[0] = eta[i + 1] / eta[i];
      a[1] = a[0] * m->csigma[i] / m->csigma[i + 1];
      for (k = 0; k < 2; k++) {
        w[k] = w[k] * a[k];
        q[i][k] = (1.0 + w[k]) / (1.0 - w[k]);
      }
    }
  }
}

/* Calc_Apq (:230-266); arrays 1-based in layer index */
static void calc_apq(const or_model *m, double lms, cplx (*Arr)[2],
                     cplx (*p)[2], cplx (*q)[2], cplx *eta,
                     cplx (*A)[OR_MAXL + 1][2]) {
  int Nl = m->Nl, i, j, k;
  cplx el[OR_MAXL + 1], w[OR_MAXL + 1][2];
  for (i = 1; i <= Nl; i++) eta[i] = csqrt(lms * lms - m->k2[i]);
  for (i = 1; i <= Nl - 1; i++) el[i] = cexp(-2.0 * eta[i] * m->thick[i]);
  el[Nl] = 0;
  calc_p(m, eta, el, lms, p);
  calc_q(m, eta, el, q);
  for (i = 1; i <= Nl; i++)
    for (k = 0; k < 2; k++) Arr[i][k] = 1.0 / eta[i] / (1.0 - p[i][k] * q[i][k] * el[i]);
  if (A) {
    for (i = 1; i <= Nl - 1; i++) {
      cplx tmp = (m->k2[i] - m->k2[i + 1]);
      tmp = m->depth[i] * tmp;
      tmp = tmp / (eta[i + 1] + eta[i]);
      for (k = 0; k < 2; k++)
        w[i][k] = (1.0 + p[i + 1][k]) / (1.0 + p[i][k] * el[i]) * cexp(tmp);
    }
    for (i = 1; i <= Nl; i++) {
      A[i][i][0] = Arr[i][0];
      A[i][i][1] = Arr[i][1];
      for (j = i - 1; j >= 1; j--) {
        for (k = 0; k < 2; k++) A[j][i][k] = A[j + 1][i][k] * w[j][k];
        A[i][j][0] = A[j][i][0];
        A[i][j][1] = A[j][i][1] * m->csigma[i] / m->csigma[j];
      }
    }
  }
}

void or_calc_apq(const or_model *m, double lm, double *Arr_o, double *p_o,
                 double *q_o, double *eta_o, double *A_o) {
  cplx Arr[OR_MAXL + 1][2], p[OR_MAXL + 1][2], q[OR_MAXL + 1][2],
      eta[OR_MAXL + 1];
  cplx(*A)[OR_MAXL + 1][2] = NULL;
  int Nl = m->Nl, i, j, k;
  if (A_o) A = calloc(OR_MAXL + 1, sizeof(*A));
  calc_apq(m, lm, Arr, p, q, eta, A);
  /* outputs in Fortran order: Arr(Nl,2), p(Nl,2), q(Nl,2), eta(Nl), A(Nl,Nl,2) */
  for (k = 0; k < 2; k++)
    for (i = 1; i <= Nl; i++) {
      ((cplx *)Arr_o)[(i - 1) + Nl * k] = Arr[i][k];
      ((cplx *)p_o)[(i - 1) + Nl * k] = p[i][k];
      ((cplx *)q_o)[(i - 1) + Nl * k] = q[i][k];
    }
  for (i = 1; i <= Nl; i++) ((cplx *)eta_o)[i - 1] = eta[i];
  if (A) {
    for (k = 0; k < 2; k++)
      for (j = 1; j <= Nl; j++)
        for (i = 1; i <= Nl; i++)
          ((cplx *)A_o)[(i - 1) + Nl * ((j - 1) + Nl * k)] = A[i][j][k];
    free(A);
  }
}

/* Calc_Double_Integral_U (ie_kernel_hankel_module.F90:44-122) with
 * Calc_pexpz_All (apq:112-137), Calc_tframe_p (83-110), Calc_qexpz_All
 * (155-177), Calc_bframe_q (138-153), Calc_FtFb (31-79), GoByShellsU
 * (ie_k:125-173), GoByShellsL (175-197).  G is complex (6,Nz,Nz) Fortran
 * order, components EXX,EXY,EXZ,EYY,EYZ,EZZ (const_module.F90:100-108). */
enum { EXX = 0, EXY = 1, EXZ = 2, EYY = 3, EYZ = 4, EZZ = 5 };
#define GIDX(c, i, j) ((c) + 6 * ((i) + Nz * (j))) /* i,j 0-based */

static void double_integral_U(const or_model *m, double lms, const double *WT,
                              cplx *G) {
  int Nz = m->Nz, Nl = m->Nl, I, l, k, Iz, Izr;
  cplx Arr[OR_MAXL + 1][2], p[OR_MAXL + 1][2], q[OR_MAXL + 1][2],
      eta[OR_MAXL + 1];
  cplx *expz = malloc(sizeof(cplx) * (Nz + 1));
  cplx(*pez)[2] = malloc(sizeof(cplx) * 2 * (Nz + 1));
  cplx(*qez)[2] = malloc(sizeof(cplx) * 2 * (Nz + 1));
  cplx(*fcont_t)[2] = malloc(sizeof(cplx) * 2 * (Nz + 1));
  cplx(*t_frame)[3] = malloc(sizeof(cplx) * 3 * (Nz + 1));
  cplx *fcont_b = malloc(sizeof(cplx) * (Nz + 1));
  cplx *b_frame = malloc(sizeof(cplx) * (Nz + 1));
  cplx(*Ft)[3] = malloc(sizeof(cplx) * 3 * (Nz + 1));
  cplx *Fb = malloc(sizeof(cplx) * (Nz + 1));
  cplx(*GII)[4] = malloc(sizeof(cplx) * 4 * (Nz + 1));
  cplx e;
  double lm2;
  (void)Nl;
  calc_apq(m, lms, Arr, p, q, eta, NULL);
  for (Iz = 1; Iz <= Nz; Iz++) {
    l = m->lnum[Iz];
    expz[Iz] = cexp(-eta[l] * (m->dz[Iz]));
  }
  /* Calc_pexpz_All */
  l = m->lnum[0];
  if (l > 1)
    e = cexp(-2.0 * eta[l] * (m->z[0] - m->depth[l - 1]));
  else if (l == 1)
    e = cexp(-2.0 * eta[1] * (m->z[0]));
  else
    e = 1.0;
  for (I = 1; I <= Nz; I++) {
    if (l != m->lnum[I]) e = 1.0;
    l = m->lnum[I];
    for (k = 0; k < 2; k++) pez[I][k] = e * p[l][k];
    e = e * expz[I] * expz[I];
  }
  /* Calc_tframe_p */
  for (I = 1; I <= Nz - 1; I++) {
    cplx a[2], b[2], c[2], w;
    l = m->lnum[I];
    for (k = 0; k < 2; k++) {
      a[k] = pez[I][k];
      b[k] = a[k] * expz[I];
      c[k] = b[k] * expz[I];
      a[k] = 1.0 + a[k];
      b[k] = 1.0 + b[k];
      c[k] = 1.0 + c[k];
    }
    w = (1.0 - expz[I]) / eta[l];
    for (k = 0; k < 2; k++) fcont_t[I][k] = expz[I] * a[k] / c[k];
    t_frame[I][0] = w * b[0] / c[0];
    t_frame[I][1] = 1.0 - fcont_t[I][1];
    t_frame[I][2] = w * b[1] / c[1];
    t_frame[I][1] = t_frame[I][1] / m->csigma[l];
    t_frame[I][2] = t_frame[I][2] / m->csigma[l];
  }
  /* Calc_qexpz_All */
  l = m->lnum[Nz];
  if (l != m->Nl)
    e = cexp(-2.0 * eta[l] * (m->depth[l] - m->z[Nz]));
  else
    e = 0.0;
  for (I = Nz; I >= 1; I--) {
    if (l != m->lnum[I]) e = 1.0;
    l = m->lnum[I];
    for (k = 0; k < 2; k++) qez[I][k] = e * q[l][k];
    e = e * expz[I] * expz[I];
  }
  /* Calc_bframe_q */
  for (I = Nz; I >= 2; I--) {
    cplx a, d;
    l = m->lnum[I];
    a = qez[I][1];
    d = 1.0 + a * expz[I] * expz[I];
    a = 1.0 + a;
    fcont_b[I] = expz[I] * a / d;
    b_frame[I] = (fcont_b[I] - 1.0) / m->csigma[l];
  }
  /* Calc_FtFb */
  for (I = 1; I <= Nz; I++) {
    cplx a[2], a1[2], a2[2], eta2, p1[2], p2[2], pe1[2], pe2[2], q1[2], q2[2],
        qe1[2], qe2[2], Fb3[3], w[2];
    l = m->lnum[I];
    for (k = 0; k < 2; k++) {
      a[k] = (1.0 - expz[I]) * Arr[l][k];
      a1[k] = a[k] / eta[l];
      a2[k] = a1[k] / eta[l];
      p1[k] = 1.0 + pez[I][k];
      p2[k] = 1.0 - pez[I][k];
      pe1[k] = 1.0 + pez[I][k] * expz[I];
      pe2[k] = -1.0 + pez[I][k] * expz[I];
      q1[k] = 1.0 + qez[I][k];
      q2[k] = 1.0 - qez[I][k];
      qe1[k] = 1.0 + qez[I][k] * expz[I];
      qe2[k] = 1.0 - qez[I][k] * expz[I];
    }
    eta2 = eta[l] * eta[l];
    for (k = 0; k < 2; k++) {
      Ft[I][k] = a1[k] * p1[k] * qe1[k];
      Fb3[k] = a1[k] * pe1[k] * q1[k];
      w[k] = -(a2[k] * (p2[k] * qe1[k] + pe1[k] * q2[k]));
    }
    Ft[I][2] = a[1] * p1[1] * qe2[1];
    Fb3[2] = a[1] * pe2[1] * q1[1];
    GII[I][0] = w[0] + 2.0 / eta2 * m->dz[I];
    GII[I][1] = Fb3[1] - Ft[I][1];
    GII[I][2] = (w[1] + 2.0 / eta2 * m->dz[I]) * m->iwm + eta2 / m->csigma[l] * w[1];
    GII[I][3] = Fb3[2] - Ft[I][2] - eta2 * w[0];
    Fb[I] = Fb3[1];
  }
  lm2 = lms * lms;
  for (Iz = 1; Iz <= Nz; Iz++) {
    cplx f1wdxx, f1wdyy, f1wdxy, f2wdx, f2wdy, f2wdzz, f3wdxx, f3wdyy, f3wdxy;
    cplx sg;
    int iz = Iz - 1;
    l = m->lnum[Iz];
    sg = m->csigma[l];
    f1wdxx = GII[Iz][0] * WT[IE_DXX] * lm2;
    f1wdyy = GII[Iz][0] * WT[IE_DYY] * lm2;
    f1wdxy = GII[Iz][0] * WT[IE_DXY] * lm2;
    f2wdx = GII[Iz][1] * WT[IE_DX];
    f2wdy = GII[Iz][1] * WT[IE_DY];
    f2wdzz = (GII[Iz][2]) * WT[IE_D0];
    f3wdxx = GII[Iz][3] * WT[IE_DXX];
    f3wdyy = GII[Iz][3] * WT[IE_DYY];
    f3wdxy = GII[Iz][3] * WT[IE_DXY];
    G[GIDX(EXX, iz, iz)] = (GII[Iz][0] * WT[IE_D0] * m->iwm + (f1wdxx + f3wdxx) / sg) + G[GIDX(EXX, iz, iz)];
    G[GIDX(EXY, iz, iz)] = (f1wdxy + f3wdxy) / sg + G[GIDX(EXY, iz, iz)];
    G[GIDX(EXZ, iz, iz)] = f2wdx / sg + G[GIDX(EXZ, iz, iz)];
    G[GIDX(EYY, iz, iz)] = (GII[Iz][0] * WT[IE_D0] * m->iwm + (f1wdyy + f3wdyy) / sg) + G[GIDX(EYY, iz, iz)];
    G[GIDX(EYZ, iz, iz)] = f2wdy / sg + G[GIDX(EYZ, iz, iz)];
    G[GIDX(EZZ, iz, iz)] = f2wdzz + G[GIDX(EZZ, iz, iz)];
    /* GoByShellsU */
    {
      cplx *fb = Ft[Iz];
      cplx u1xx = fb[0] * (WT[IE_DXX] + WT[IE_D0]) * m->iwm;
      cplx u1yy = fb[0] * (WT[IE_DYY] + WT[IE_D0]) * m->iwm;
      cplx u1xy = fb[0] * WT[IE_DXY] * m->iwm;
      cplx u2x = fb[1] * WT[IE_DX], u2y = fb[1] * WT[IE_DY];
      cplx u2zz = fb[1] * WT[IE_D0] * lm2;
      cplx u3xx = fb[2] * WT[IE_DXX], u3yy = fb[2] * WT[IE_DYY], u3xy = fb[2] * WT[IE_DXY];
      for (Izr = Iz - 1; Izr >= 1; Izr--) {
        int ir = Izr - 1;
        cplx *fr = t_frame[Izr];
        G[GIDX(EXX, ir, iz)] = u1xx * fr[0] + u3xx * fr[1] + G[GIDX(EXX, ir, iz)];
        G[GIDX(EXY, ir, iz)] = u3xy * fr[1] + u1xy * fr[0] + G[GIDX(EXY, ir, iz)];
        G[GIDX(EXZ, ir, iz)] = u2x * fr[1] + G[GIDX(EXZ, ir, iz)];
        G[GIDX(EYY, ir, iz)] = u1yy * fr[0] + u3yy * fr[1] + G[GIDX(EYY, ir, iz)];
        G[GIDX(EYZ, ir, iz)] = u2y * fr[1] + G[GIDX(EYZ, ir, iz)];
        G[GIDX(EZZ, ir, iz)] = u2zz * fr[2] + G[GIDX(EZZ, ir, iz)];
        u1xx = u1xx * fcont_t[Izr][0];
        u1yy = u1yy * fcont_t[Izr][0];
        u1xy = u1xy * fcont_t[Izr][0];
        u2x = u2x * fcont_t[Izr][1];
        u2y = u2y * fcont_t[Izr][1];
        u2zz = u2zz * fcont_t[Izr][1];
        u3xx = u3xx * fcont_t[Izr][1];
        u3yy = u3yy * fcont_t[Izr][1];
        u3xy = u3xy * fcont_t[Izr][1];
      }
    }
    /* GoByShellsL */
    {
      cplx v2x = Fb[Iz] * WT[IE_DX], v2y = Fb[Iz] * WT[IE_DY];
      for (Izr = Iz + 1; Izr <= Nz; Izr++) {
        int ir = Izr - 1;
        G[GIDX(EXZ, ir, iz)] = v2x * b_frame[Izr] + G[GIDX(EXZ, ir, iz)];
        G[GIDX(EYZ, ir, iz)] = v2y * b_frame[Izr] + G[GIDX(EYZ, ir, iz)];
        v2x = v2x * fcont_b[Izr];
        v2y = v2y * fcont_b[Izr];
      }
    }
  }
  free(expz);
  free(pez);
  free(qez);
  free(fcont_t);
  free(t_frame);
  free(fcont_b);
  free(b_frame);
  free(Ft);
  free(Fb);
  free(GII);
}

void or_double_integral_U(const or_model *m, double lm, const double *W,
                          double *G) {
  double_integral_U(m, lm, W, (cplx *)G);
}

/* ------------------------------------------------------------------------ */
/* Real-space tensor: calc_ie_tensor_module.F90                              */
/* ------------------------------------------------------------------------ */
int64_t or_block_size(int Nz) {
  return 4 * (int64_t)Nz * (Nz + 1) / 2 + 2 * (int64_t)Nz * Nz;
}

static const double WT_THRESHOLD = 1e-12; /* :37 */

/* CalcIntegralGreenTensor3dElementGeneral (:241-293) */
void or_tensor_element(const or_model *m, int lx, int ly, double *blk_d) {
  int Nz = m->Nz, P = Nz * (Nz + 1) / 2, K, c, i, j;
  double rmin = sqrt(m->dx * m->dx + m->dy * m->dy) / 2.0; /* :102 */
  double lms[OR_NTAPS], WT[OR_NTAPS * 6], W0[6];
  cplx *G1 = calloc((size_t)6 * Nz * Nz, sizeof(cplx));
  cplx *blk = (cplx *)blk_d;
  or_vbt_lms(rmin, lms);
  or_compute_weights(lx, ly, m->dx, m->dy, rmin, WT, W0);
  for (K = 0; K < OR_NTAPS; K++) {
    double Wm = 0, W[6], lm;
    for (c = 0; c < 6; c++) { /* :266 Wm = ABS(WT(K,:)/W0(:)) */
      double t = fabs(WT[K + OR_NTAPS * c] / W0[c]);
      if (c == 0 || t > Wm) Wm = t;
    }
    if (Wm > WT_THRESHOLD) {
      lm = lms[K];
      W[IE_DXX] = WT[K + OR_NTAPS * IE_DXX] / lm;
      W[IE_DYY] = WT[K + OR_NTAPS * IE_DYY] / lm;
      W[IE_DXY] = WT[K + OR_NTAPS * IE_DXY] / lm;
      W[IE_DX] = WT[K + OR_NTAPS * IE_DX] * lm;
      W[IE_DY] = WT[K + OR_NTAPS * IE_DY] * lm;
      W[IE_D0] = WT[K + OR_NTAPS * IE_D0] * lm;
      double_integral_U(m, lm, W, G1);
    }
  }
  /* pack (:278-288) */
  for (j = 0; j < Nz; j++)
    for (i = 0; i < Nz; i++) {
      blk[4 * P + i + Nz * j] = G1[GIDX(EXZ, i, j)] / PI / 4.0;
      blk[4 * P + Nz * Nz + i + Nz * j] = G1[GIDX(EYZ, i, j)] / PI / 4.0;
    }
  for (j = 0; j < Nz; j++)
    for (i = 0; i <= j; i++) {
      int pidx = i + j * (j + 1) / 2;
      blk[0 * P + pidx] = G1[GIDX(EXX, i, j)] / PI / 4.0;
      blk[1 * P + pidx] = G1[GIDX(EXY, i, j)] / PI / 4.0;
      blk[2 * P + pidx] = G1[GIDX(EYY, i, j)] / PI / 4.0;
      blk[3 * P + pidx] = G1[GIDX(EZZ, i, j)] / PI / 4.0;
    }
  free(G1);
}

/* Rounding scale of or_tensor_element.  G is linear in the six weights of
 * every tap, G = sum_K sum_c WT_c(K) s_c(lambda_K) B_c(lambda_K), with the
 * scalings s_c = 1/lambda (DXX, DYY, DXY) or lambda (DX, DY, D0) of
 * calc_ie_tensor_module.F90:269-274.  Two FP64 evaluations of the weights
 * (the FFT deconvolution of VolumeBesselTransforms) differ by a few eps times
 * the largest |WT_c| of each vector, at every tap alike (at taps of tiny
 * lambda that is O(1) relative), so per packed entry this returns
 *   S = sum over unmasked taps K and over c of
 *       max_K' |WT_c(K')| |s_c(lambda_K)| |B_c(lambda_K)| / (4 pi):
 * weight errors of delta max|WT_c| change G by at most delta S, however much
 * the far-offset tap sums cancel. */
void or_tensor_element_abssum(const or_model *m, int lx, int ly, double *out) {
  int Nz = m->Nz, P = Nz * (Nz + 1) / 2, K, c, i, j;
  double rmin = sqrt(m->dx * m->dx + m->dy * m->dy) / 2.0;
  double lms[OR_NTAPS], WT[OR_NTAPS * 6], W0[6], Ws[OR_NTAPS * 6], Wmax[6] = {0};
  int keep[OR_NTAPS];
  cplx *G1 = malloc((size_t)6 * Nz * Nz * sizeof(cplx));
  double *A = calloc((size_t)6 * Nz * Nz, sizeof(double));
  or_vbt_lms(rmin, lms);
  or_compute_weights(lx, ly, m->dx, m->dy, rmin, WT, W0);
  for (K = 0; K < OR_NTAPS; K++) {
    double Wm = 0, lm = lms[K];
    for (c = 0; c < 6; c++) {
      double t = fabs(WT[K + OR_NTAPS * c] / W0[c]);
      if (c == 0 || t > Wm) Wm = t;
    }
    keep[K] = Wm > WT_THRESHOLD;
    /* |scaling| of each weight at this tap */
    Ws[6 * K + IE_DXX] = Ws[6 * K + IE_DYY] = Ws[6 * K + IE_DXY] = 1.0 / lm;
    Ws[6 * K + IE_DX] = Ws[6 * K + IE_DY] = Ws[6 * K + IE_D0] = lm;
    for (c = 0; c < 6; c++)
      if (fabs(WT[K + OR_NTAPS * c]) > Wmax[c]) Wmax[c] = fabs(WT[K + OR_NTAPS * c]);
  }
  for (K = 0; K < OR_NTAPS; K++) {
    if (!keep[K]) continue;
    for (c = 0; c < 6; c++) {
      double W[6] = {0, 0, 0, 0, 0, 0};
      int64_t e;
      if (Wmax[c] == 0) continue;
      W[c] = 1.0;
      memset(G1, 0, (size_t)6 * Nz * Nz * sizeof(cplx));
      double_integral_U(m, lms[K], W, G1);
      for (e = 0; e < (int64_t)6 * Nz * Nz; e++) A[e] += Wmax[c] * Ws[6 * K + c] * cabs(G1[e]);
    }
  }
  for (j = 0; j < Nz; j++)
    for (i = 0; i < Nz; i++) {
      out[4 * P + i + Nz * j] = A[GIDX(EXZ, i, j)] / PI / 4.0;
      out[4 * P + Nz * Nz + i + Nz * j] = A[GIDX(EYZ, i, j)] / PI / 4.0;
    }
  for (j = 0; j < Nz; j++)
    for (i = 0; i <= j; i++) {
      int pidx = i + j * (j + 1) / 2;
      out[0 * P + pidx] = A[GIDX(EXX, i, j)] / PI / 4.0;
      out[1 * P + pidx] = A[GIDX(EXY, i, j)] / PI / 4.0;
      out[2 * P + pidx] = A[GIDX(EYY, i, j)] / PI / 4.0;
      out[3 * P + pidx] = A[GIDX(EZZ, i, j)] / PI / 4.0;
    }
  free(G1);
  free(A);
}

/* CalcIntegralGreenTensor (:54-192), single rank: every (lx,ly) offset */
void or_tensor_realspace(const or_model *m, double *Greal, int nthreads) {
  int Nx = m->Nx, Ny = m->Ny;
  int64_t BS = or_block_size(m->Nz), b;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (b = 0; b < (int64_t)Nx * Ny; b++) {
    int lx = (int)(b % Nx), ly = (int)(b / Nx);
    or_tensor_element(m, lx, ly, Greal + 2 * BS * b);
  }
}

/* ------------------------------------------------------------------------ */
/* Tensor spectrum: CalcFFTofIETensorGeneral / CALC_FFT_OF_TENSOR_COMPONENT  */
/* (Integral_Equation_Module.F90:319-471) with DistributedFourierYMirror     */
/* (Distributed_FFT.F90:507-551).                                           */
/* ------------------------------------------------------------------------ */
/* extension signs per block entry: xx(+,+) xy(-,-) yy(+,+) zz(+,+) xz(-,+) yz(+,-) */
static void entry_signs(int Nz, int64_t e, int *sx, int *sy) {
  int64_t P = (int64_t)Nz * (Nz + 1) / 2;
  if (e < P) { *sx = 1; *sy = 1; }
  else if (e < 2 * P) { *sx = -1; *sy = -1; }
  else if (e < 4 * P) { *sx = 1; *sy = 1; }
  else if (e < 4 * P + (int64_t)Nz * Nz) { *sx = -1; *sy = 1; }
  else { *sx = 1; *sy = -1; }
}

void or_tensor_fft(const or_model *m, const double *Greal_d, double *Gspec_d,
                   int nthreads) {
  int Nx = m->Nx, Ny = m->Ny, Nx2 = 2 * Nx, Ny2 = 2 * Ny;
  int64_t BS = or_block_size(m->Nz), e;
  const cplx *Greal = (const cplx *)Greal_d;
  cplx *Gspec = (cplx *)Gspec_d;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (e = 0; e < BS; e++) {
    cplx *F = calloc((size_t)Nx2 * Ny2, sizeof(cplx));
    cplx *col = malloc(sizeof(cplx) * Ny2);
    int sx, sy, lx, ly, kx, ky;
    entry_signs(m->Nz, e, &sx, &sy);
    /* symmetric extension, offset N set to zero (:441-447, Distributed_FFT:526-531) */
    for (ly = 0; ly < Ny; ly++)
      for (lx = 0; lx < Nx; lx++) {
        cplx g = Greal[e + BS * (lx + (int64_t)Nx * ly)];
        F[lx + Nx2 * ly] = g;
        if (lx > 0) F[(Nx2 - lx) + Nx2 * ly] = sx * g;
        if (ly > 0) F[lx + Nx2 * (Ny2 - ly)] = sy * g;
        if (lx > 0 && ly > 0) F[(Nx2 - lx) + Nx2 * (Ny2 - ly)] = sx * sy * g;
      }
    /* forward 2-D DFT (FFTW_FORWARD = exp(-i)) */
    for (ly = 0; ly < Ny2; ly++) fft_c(F + Nx2 * ly, Nx2, -1);
    for (kx = 0; kx < Nx2; kx++) {
      for (ly = 0; ly < Ny2; ly++) col[ly] = F[kx + Nx2 * ly];
      fft_c(col, Ny2, -1);
      /* y-Nyquist cancellation (Distributed_FFT.F90:540-545) */
      {
        cplx t = col[Ny];
        double s = 1.0;
        for (ky = 0; ky < Ny2; ky++) {
          col[ky] = col[ky] - s * t;
          s = -s;
        }
      }
      for (ly = 0; ly < Ny2; ly++) F[kx + Nx2 * ly] = col[ly];
    }
    /* x-Nyquist cancellation with ss = +1 on the kx = 0.. half
     * (Integral_Equation_Module.F90:459-466) */
    for (ky = 0; ky < Ny; ky++) {
      cplx t = F[Nx + Nx2 * ky];
      double s = 1.0;
      for (kx = 0; kx < Nx; kx++) {
        Gspec[e + BS * (kx + (int64_t)Nx * ky)] = F[kx + Nx2 * ky] - t * s;
        s = -s;
      }
    }
    free(col);
    free(F);
  }
}

/* ------------------------------------------------------------------------ */
/* Operator apply: APPLY_PRECOND / MULT_IE_OP / VERTICAL_MULT_GENERAL        */
/* ------------------------------------------------------------------------ */
static double g_apply_t[3];
void or_apply_timings(double *t3) {
  t3[0] = g_apply_t[0];
  t3[1] = g_apply_t[1];
  t3[2] = g_apply_t[2];
}

/* sqsigb, csigb per layer (PreparePreconditioner, calc_ie_tensor_module.F90:295-306) */
static void layer_sig(const or_model *m, int iz /*0-based*/, cplx *csigb,
                      double *sqsigb) {
  *csigb = m->csigma[m->lnum[iz + 1]];
  *sqsigb = sqrt(creal(*csigb));
}

/* ZSPMV('U') semantic: y += alpha * S x, S complex-symmetric packed upper */
static void spmv_u(int n, cplx alpha, const cplx *S, const cplx *x, cplx *y) {
  int i, j;
  for (i = 0; i < n; i++) {
    cplx acc = 0;
    for (j = 0; j < n; j++) {
      int lo = i < j ? i : j, hi = i < j ? j : i;
      acc += S[lo + hi * (hi + 1) / 2] * x[j];
    }
    y[i] += alpha * acc;
  }
}
/* ZGEMV('N'/'T') semantic: y += alpha * op(A) x, A column-major n x n */
static void gemv(int trans, int n, cplx alpha, const cplx *A, const cplx *x,
                 cplx *y) {
  int i, j;
  for (i = 0; i < n; i++) {
    cplx acc = 0;
    for (j = 0; j < n; j++) acc += (trans ? A[j + n * i] : A[i + n * j]) * x[j];
    y[i] += alpha * acc;
  }
}

/* One (kx or -kx, ky or -ky) point: out = G(sigma, mirror) in
 * (VERTICAL_MULT_GENERAL, apply_ie_operator_module.F90:276-292):
 *   out_x = Sxx x + s m Sxy y + s Axz z
 *   out_y = s m Sxy x + Syy y + m Ayz z
 *   out_z = -s Axz^T x - m Ayz^T y + Szz z                                  */
static void point_mult(int Nz, const cplx *blk, double s, double mir,
                       const cplx *in, cplx *out) {
  int P = Nz * (Nz + 1) / 2;
  const cplx *Sxx = blk, *Sxy = blk + P, *Syy = blk + 2 * P, *Szz = blk + 3 * P;
  const cplx *Axz = blk + 4 * P, *Ayz = blk + 4 * P + Nz * Nz;
  const cplx *ix = in, *iy = in + Nz, *iz = in + 2 * Nz;
  cplx *ox = out, *oy = out + Nz, *oz = out + 2 * Nz;
  memset(out, 0, sizeof(cplx) * 3 * Nz);
  spmv_u(Nz, 1.0, Sxx, ix, ox);
  spmv_u(Nz, s * mir, Sxy, iy, ox);
  spmv_u(Nz, s * mir, Sxy, ix, oy);
  spmv_u(Nz, 1.0, Syy, iy, oy);
  spmv_u(Nz, 1.0, Szz, iz, oz);
  gemv(0, Nz, s, Axz, iz, ox);
  gemv(1, Nz, -s, Axz, ix, oz);
  gemv(0, Nz, mir, Ayz, iz, oy);
  gemv(1, Nz, -mir, Ayz, iy, oz);
}

void or_apply(const or_model *m, const double *Gspec_d, const double *csiga_d,
              const double *vin_d, double *vout_d, int nthreads) {
  int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz, Nx2 = 2 * Nx, Ny2 = 2 * Ny;
  int64_t BS = or_block_size(Nz), NS = (int64_t)Nx2 * Ny2, s3, b;
  const cplx *Gspec = (const cplx *)Gspec_d, *csiga = (const cplx *)csiga_d,
             *vin = (const cplx *)vin_d;
  cplx *vout = (cplx *)vout_d;
  cplx *buf = calloc((size_t)NS * 3 * Nz, sizeof(cplx)); /* [cz][ky][kx] */
  double t0, t1, t2, t3;
  double Nnorm = 4.0 * Nx * Ny; /* :56 */
  if (nthreads < 1) nthreads = 1;
  t0 = now_s();
  /* pre-scale into the zero-padded buffer (:57-115) and forward 2-D FFT */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (s3 = 0; s3 < 3 * Nz; s3++) {
    int c = (int)(s3 / Nz), z = (int)(s3 % Nz), x, y, kx;
    cplx *F = buf + NS * s3, csigb, *col;
    double sq;
    layer_sig(m, z, &csigb, &sq);
    for (y = 0; y < Ny; y++)
      for (x = 0; x < Nx; x++) {
        cplx sa = csiga[x + (int64_t)Nx * (y + (int64_t)Ny * z)];
        cplx asiga = 2.0 * sq / (sa + conj(csigb));
        cplx gsig = (sa - csigb) * asiga;
        F[x + Nx2 * y] = vin[x + (int64_t)Nx * (y + (int64_t)Ny * (z + (int64_t)Nz * c))] * gsig;
      }
    for (y = 0; y < Ny; y++) fft_c(F + Nx2 * y, Nx2, -1);
    col = malloc(sizeof(cplx) * Ny2);
    for (kx = 0; kx < Nx2; kx++) {
      for (y = 0; y < Ny2; y++) col[y] = F[kx + Nx2 * y];
      fft_c(col, Ny2, -1);
      for (y = 0; y < Ny2; y++) F[kx + Nx2 * y] = col[y];
    }
    free(col);
  }
  t1 = now_s();
  /* contraction per (kx-pair, ky) with the four (+-kx, +-ky) points; the
   * x- and y-Nyquist lines are dropped (Distributed_FFT.F90:464-474,723-724) */
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (b = 0; b < (int64_t)Nx * Ny; b++) {
    int kx = (int)(b % Nx), ky = (int)(b / Nx), p, i;
    const cplx *blk = Gspec + BS * b;
    cplx *in = malloc(sizeof(cplx) * 3 * Nz), *out = malloc(sizeof(cplx) * 3 * Nz);
    for (p = 0; p < 4; p++) {
      int sgx = (p & 1) ? -1 : 1, sgy = (p & 2) ? -1 : 1;
      int px = sgx > 0 ? kx : (kx == 0 ? Nx : Nx2 - kx);
      int py = sgy > 0 ? ky : (ky == 0 ? Ny : Ny2 - ky);
      int64_t off = px + (int64_t)Nx2 * py;
      if ((sgx < 0 && kx == 0) || (sgy < 0 && ky == 0)) {
        for (i = 0; i < 3 * Nz; i++) buf[off + NS * i] = 0; /* Nyquist */
        continue;
      }
      for (i = 0; i < 3 * Nz; i++) in[i] = buf[off + NS * i];
      point_mult(Nz, blk, (double)sgx, (double)sgy, in, out);
      for (i = 0; i < 3 * Nz; i++) buf[off + NS * i] = out[i];
    }
    free(in);
    free(out);
  }
  t2 = now_s();
  /* backward 2-D FFT and post-scale (:125-174) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (s3 = 0; s3 < 3 * Nz; s3++) {
    int c = (int)(s3 / Nz), z = (int)(s3 % Nz), x, y, kx;
    cplx *F = buf + NS * s3, csigb, *col;
    double sq, dzz = m->dz[z + 1];
    layer_sig(m, z, &csigb, &sq);
    col = malloc(sizeof(cplx) * Ny2);
    for (kx = 0; kx < Nx2; kx++) {
      for (y = 0; y < Ny2; y++) col[y] = F[kx + Nx2 * y];
      fft_c(col, Ny2, +1);
      for (y = 0; y < Ny2; y++) F[kx + Nx2 * y] = col[y];
    }
    free(col);
    for (y = 0; y < Ny; y++) fft_c(F + Nx2 * y, Nx2, +1);
    for (y = 0; y < Ny; y++)
      for (x = 0; x < Nx; x++) {
        int64_t vi = x + (int64_t)Nx * (y + (int64_t)Ny * (z + (int64_t)Nz * c));
        cplx sa = csiga[x + (int64_t)Nx * (y + (int64_t)Ny * z)];
        cplx asiga = 2.0 * sq / (sa + conj(csigb));
        cplx d1 = vin[vi] * asiga;
        cplx d2 = d1 - F[x + Nx2 * y] / dzz / Nnorm;
        vout[vi] = d2 * sq;
      }
  }
  t3 = now_s();
  g_apply_t[0] = t1 - t0;
  g_apply_t[1] = t2 - t1;
  g_apply_t[2] = t3 - t2;
  free(buf);
}

/* Direct block-Toeplitz sum: the exact zero-padded linear convolution that
 * the FFT path computes, with no FFT and no Nyquist handling.  Real-space
 * kernel K_cc'(lx,ly;z,z') extended to negative offsets with the signs of
 * Integral_Equation_Module.F90:334-346. */
void or_apply_bruteforce(const or_model *m, const double *Greal_d,
                         const double *csiga_d, const double *vin_d,
                         double *vout_d, int nthreads) {
  int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz, P = Nz * (Nz + 1) / 2;
  int64_t BS = or_block_size(Nz), nc = (int64_t)Nx * Ny * Nz, r;
  const cplx *Greal = (const cplx *)Greal_d, *csiga = (const cplx *)csiga_d,
             *vin = (const cplx *)vin_d;
  cplx *vout = (cplx *)vout_d;
  cplx *gv = malloc(sizeof(cplx) * 3 * nc);
  int x, y, z, c;
  if (nthreads < 1) nthreads = 1;
  for (c = 0; c < 3; c++)
    for (z = 0; z < Nz; z++)
      for (y = 0; y < Ny; y++)
        for (x = 0; x < Nx; x++) {
          cplx csigb, sa;
          double sq;
          int64_t ci = x + (int64_t)Nx * (y + (int64_t)Ny * z);
          layer_sig(m, z, &csigb, &sq);
          sa = csiga[ci];
          gv[ci + nc * c] = vin[ci + nc * c] * ((sa - csigb) * (2.0 * sq / (sa + conj(csigb))));
        }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (r = 0; r < (int64_t)Nx * Ny; r++) {
    int xr = (int)(r % Nx), yr = (int)(r / Nx), xs, ys, zr, zs;
    cplx acc[3 * 64];
    for (zr = 0; zr < Nz; zr++) acc[zr] = acc[Nz + zr] = acc[2 * Nz + zr] = 0;
    for (ys = 0; ys < Ny; ys++)
      for (xs = 0; xs < Nx; xs++) {
        int lx = xr - xs, ly = yr - ys;
        int ax = lx < 0 ? -lx : lx, ay = ly < 0 ? -ly : ly;
        double fx = lx < 0 ? -1.0 : 1.0, fy = ly < 0 ? -1.0 : 1.0;
        const cplx *blk = Greal + BS * (ax + (int64_t)Nx * ay);
        for (zr = 0; zr < Nz; zr++)
          for (zs = 0; zs < Nz; zs++) {
            int lo = zr < zs ? zr : zs, hi = zr < zs ? zs : zr;
            int pi = lo + hi * (hi + 1) / 2;
            int64_t si = xs + (int64_t)Nx * (ys + (int64_t)Ny * zs);
            cplx vx = gv[si], vy = gv[si + nc], vz = gv[si + 2 * nc];
            cplx Sxx = blk[pi], Sxy = blk[P + pi] * (fx * fy), Syy = blk[2 * P + pi],
                 Szz = blk[3 * P + pi];
            cplx Axz = blk[4 * P + zr + Nz * zs] * fx;          /* xz(+-,+) */
            cplx Azx = -blk[4 * P + zs + Nz * zr] * fx;         /* zx = -xz^T */
            cplx Ayz = blk[4 * P + Nz * Nz + zr + Nz * zs] * fy; /* yz(+,-) */
            cplx Azy = -blk[4 * P + Nz * Nz + zs + Nz * zr] * fy;
            acc[zr] += Sxx * vx + Sxy * vy + Axz * vz;
            acc[Nz + zr] += Sxy * vx + Syy * vy + Ayz * vz;
            acc[2 * Nz + zr] += Azx * vx + Azy * vy + Szz * vz;
          }
      }
    for (c = 0; c < 3; c++)
      for (zr = 0; zr < Nz; zr++) {
        cplx csigb, sa, asiga;
        double sq;
        int64_t ci = xr + (int64_t)Nx * (yr + (int64_t)Ny * zr);
        layer_sig(m, zr, &csigb, &sq);
        sa = csiga[ci];
        asiga = 2.0 * sq / (sa + conj(csigb));
        vout[ci + nc * c] = (vin[ci + nc * c] * asiga - acc[c * Nz + zr] / m->dz[zr + 1]) * sq;
      }
  }
  free(gv);
}

/* ------------------------------------------------------------------------ */
/* Plane-wave RHS: PlaneWaveIntegral (sources_module.F90:67-101)            */
/* ------------------------------------------------------------------------ */
void or_plane_wave_integral(const or_model *m, int pol, double *E0_d) {
  int Nl = m->Nl, Nx = m->Nx, Ny = m->Ny, Nz = m->Nz, Iz, l;
  int64_t nc = (int64_t)Nx * Ny * Nz, i;
  cplx Arr[OR_MAXL + 1][2], p[OR_MAXL + 1][2], q[OR_MAXL + 1][2],
      eta[OR_MAXL + 1], t1;
  cplx(*A)[OR_MAXL + 1][2] = calloc(OR_MAXL + 1, sizeof(*A));
  cplx *E0 = (cplx *)E0_d;
  memset(E0, 0, sizeof(cplx) * 3 * nc);
  calc_apq(m, 0.0, Arr, p, q, eta, A);
  if (Nl > 1)
    t1 = 1.0 + q[1][0] * cexp(-m->depth[1] * eta[1] * 2.0);
  else
    t1 = 1.0;
  t1 = A[1][1][0] * t1;
  for (Iz = 1; Iz <= Nz; Iz++) {
    double dz = m->z[Iz] - m->z[Iz - 1];
    cplx e1, e2, e3, e4, G;
    l = m->lnum[Iz];
    e1 = cexp(-eta[l] * m->z[Iz - 1]);
    e2 = cexp(-eta[l] * m->z[Iz]);
    if (l < Nl) {
      e3 = q[l][0] * (cexp(-eta[l] * (2.0 * m->depth[l] - m->z[Iz - 1])));
      e4 = q[l][0] * (cexp(-eta[l] * (2.0 * m->depth[l] - m->z[Iz])));
    } else {
      e3 = e4 = 0;
    }
    G = (e1 - e2 + e4 - e3) * A[l][1][0] / t1;
    G = G / dz / eta[l];
    for (i = 0; i < (int64_t)Nx * Ny; i++) E0[i + (int64_t)Nx * Ny * (Iz - 1) + nc * pol] = G;
  }
  free(A);
}

/* ------------------------------------------------------------------------ */
/* GFGMRES (GFGMRES/gmres.F90, seq_fgmres.F90) -- depth-2: flexible outer   */
/* FGMRES with one cycle of plain GMRES as the right preconditioner.         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t N;
  int M;
  int flex;
  int maxit;
  double tol;
  cplx *x, *y, *r0, *w, *dp, *V, *H, *Z, *gs;
  double *gc;
  int iternum;
  struct gm_solver *inner;
  or_matvec_fn A;
  void *ctx;
  or_gmres_info *stats;
  double *hist;
  int hist_len;
  int is_outer;
} gm_solver;

#define HH(s, i, j) ((s)->H[(i) + (int64_t)((s)->M + 1) * (j)]) /* 0-based */

static void gm_alloc(gm_solver *s, int64_t N, int M, int flex) {
  s->N = N;
  s->M = M;
  s->flex = flex;
  s->x = malloc(sizeof(cplx) * N);
  s->r0 = malloc(sizeof(cplx) * N);
  s->w = malloc(sizeof(cplx) * N);
  s->y = malloc(sizeof(cplx) * (M + 1));
  s->dp = malloc(sizeof(cplx) * (M + 1));
  s->V = malloc(sizeof(cplx) * N * (M + 1));
  s->H = calloc((size_t)(M + 1) * (M + 1), sizeof(cplx));
  s->Z = flex ? malloc(sizeof(cplx) * N * M) : NULL;
  s->gs = malloc(sizeof(cplx) * (M + 1));
  s->gc = malloc(sizeof(double) * (M + 1));
}
static void gm_free(gm_solver *s) {
  free(s->x); free(s->r0); free(s->w); free(s->y); free(s->dp);
  free(s->V); free(s->H); free(s->Z); free(s->gs); free(s->gc);
}

static void gm_apply(gm_solver *s, const cplx *in, cplx *out) {
  s->A(s->ctx, (const double *)in, (double *)out);
  s->stats->matvecs++;
}
/* DISTRIBUTED_DOT_PRODUCT (ie_solver_module.F90:299-333): res = M^H v */
static void gm_dots(gm_solver *s, const cplx *Mv, int K, const cplx *v,
                    cplx *res) {
  int k;
  for (k = 0; k < K; k++) {
    cplx acc = 0;
    int64_t i;
    const cplx *col = Mv + s->N * k;
    for (i = 0; i < s->N; i++) acc += conj(col[i]) * v[i];
    res[k] = acc;
  }
  s->stats->dotprods += K;
}
/* CalculateVectorNorm (gmres.F90:537-549) */
static double gm_norm(gm_solver *s, const cplx *v) {
  cplx t;
  gm_dots(s, v, 1, v, &t);
  return sqrt(creal(t));
}
/* reference BLAS ZROTG */
static void zrotg(cplx *ca, cplx cb, double *c, cplx *s) {
  if (cabs(*ca) == 0.0) {
    *c = 0.0;
    *s = 1.0;
    *ca = cb;
  } else {
    double scale = cabs(*ca) + cabs(cb);
    double a1 = cabs(*ca / scale), b1 = cabs(cb / scale);
    double norm = scale * sqrt(a1 * a1 + b1 * b1);
    cplx alpha = *ca / cabs(*ca);
    *c = cabs(*ca) / norm;
    *s = alpha * conj(cb) / norm;
    *ca = alpha * norm;
  }
}
/* ZROT with real c, complex s */
static void zrot(cplx *cx, cplx *cy, double c, cplx s) {
  cplx t = c * (*cx) + s * (*cy);
  *cy = c * (*cy) - conj(s) * (*cx);
  *cx = t;
}

static void gm_solve(gm_solver *s, cplx *x, const cplx *x0, const cplx *b,
                     or_gmres_info *info);

/* ApplyRightPreconditioner (gmres.F90:491-506) */
static void gm_right_precond(gm_solver *s, const cplx *vin, cplx *vout, int K) {
  if (s->inner) {
    or_gmres_info ii;
    gm_solve((gm_solver *)s->inner, vout, vin, vin, &ii); /* seq_fgmres.F90:225 */
  } else {
    memcpy(vout, vin, sizeof(cplx) * s->N);
  }
  if (s->flex) memcpy(s->Z + s->N * K, vout, sizeof(cplx) * s->N);
}

/* ConstructCurrentSolution (:423-458) */
static void gm_construct(gm_solver *s, int jH, cplx *solution) {
  int L = jH - 1, i, j;
  int64_t n;
  int M = s->M;
  for (i = 0; i < L; i++) s->y[i] = HH(s, i, M);
  for (i = L - 1; i >= 0; i--) { /* ZTRSV('U','N','N') */
    cplx t = s->y[i];
    for (j = i + 1; j < L; j++) t -= HH(s, i, j) * s->y[j];
    s->y[i] = t / HH(s, i, i);
  }
  /* CalculateSolutionProjectionToKrylovSubspace (:461-485) */
  for (n = 0; n < s->N; n++) {
    cplx acc = 0;
    const cplx *B = s->flex ? s->Z : s->V;
    for (j = 0; j < L; j++) acc += B[n + s->N * j] * s->y[j];
    s->x[n] = acc;
  }
  if (s->flex)
    memcpy(s->r0, s->x, sizeof(cplx) * s->N);
  else
    gm_right_precond(s, s->x, s->r0, L - 1 < 0 ? 0 : L - 1);
  for (n = 0; n < s->N; n++) {
    s->x[n] = s->r0[n] + solution[n];
    solution[n] = s->x[n];
  }
}

/* FinalActions (:382-421) */
static double gm_final(gm_solver *s, int jH, double sPb, const cplx *rhs,
                       cplx *solution, int *stat) {
  int64_t n;
  double tr, res;
  gm_construct(s, jH, solution);
  memcpy(s->w, solution, sizeof(cplx) * s->N);
  gm_apply(s, s->w, s->r0);
  for (n = 0; n < s->N; n++) {
    s->r0[n] = rhs[n] - s->r0[n];
    s->w[n] = s->r0[n];
  }
  tr = gm_norm(s, s->w);
  res = tr / sPb;
  *stat = (res >= s->tol) ? 2 : 1;
  return res;
}

/* KrylovSubspaceConstruction (:216-329) */
static double gm_krylov(gm_solver *s, double bn, int *jHp) {
  int M = s->M, jH = *jHp, i, northo;
  int64_t n;
  double bea = -1.0;
  for (;;) {
    int j = jH - 1; /* 0-based column */
    double dloo, dnormw = 0;
    cplx h1, h2;
    gm_right_precond(s, s->V + s->N * j, s->w, j);
    gm_apply(s, s->w, s->r0);
    memcpy(s->w, s->r0, sizeof(cplx) * s->N);
    for (i = 0; i < jH; i++) HH(s, i, j) = 0;
    northo = 0;
    for (;;) { /* iterated classical Gram-Schmidt */
      northo++;
      gm_dots(s, s->V, jH, s->w, s->dp);
      for (i = 0; i < jH; i++) HH(s, i, j) += s->dp[i];
      for (n = 0; n < s->N; n++) {
        cplx acc = 0;
        for (i = 0; i < jH; i++) acc += s->V[n + s->N * i] * s->dp[i];
        s->w[n] -= acc;
      }
      dloo = 0;
      for (i = 0; i < jH; i++) dloo += creal(s->dp[i] * conj(s->dp[i]));
      dloo = sqrt(dloo);
      dnormw = gm_norm(s, s->w);
      if ((2.0 * dnormw > dloo) || (northo > 3)) break;
    }
    HH(s, jH, j) = dnormw;
    for (n = 0; n < s->N; n++) s->V[n + s->N * jH] = s->w[n] / dnormw;
    for (i = 0; i < j; i++) zrot(&HH(s, i, j), &HH(s, i + 1, j), s->gc[i], s->gs[i]);
    h1 = HH(s, j, j);
    h2 = HH(s, j + 1, j);
    zrotg(&h1, h2, &s->gc[j], &s->gs[j]);
    zrot(&HH(s, j, M), &HH(s, j + 1, M), s->gc[j], s->gs[j]);
    zrot(&HH(s, j, j), &HH(s, j + 1, j), s->gc[j], s->gs[j]);
    HH(s, j + 1, j) = 0;
    bea = cabs(HH(s, j + 1, M)) / bn;
    s->iternum++;
    if (s->is_outer) {
      if (s->hist && s->iternum - 1 < s->hist_len) s->hist[s->iternum - 1] = bea;
    }
    jH++;
    if ((bea < s->tol) || (jH > M)) break;
  }
  *jHp = jH;
  return bea;
}

/* PrepareRestart (:333-379), iterative residual */
static double gm_restart(gm_solver *s, const cplx *rhs, cplx *solution, int jH) {
  int M = s->M, J;
  int64_t n;
  double beta;
  (void)rhs;
  gm_construct(s, jH, solution);
  beta = cabs(HH(s, M, M));
  for (J = M - 1; J >= 0; J--) {
    HH(s, J, M) = 0;
    zrot(&HH(s, J, M), &HH(s, J + 1, M), s->gc[J], -s->gs[J]);
  }
  for (n = 0; n < s->N; n++) {
    cplx acc = 0;
    for (J = 0; J <= M; J++) acc += s->V[n + s->N * J] * HH(s, J, M);
    s->w[n] = acc;
  }
  return beta;
}

/* GMRES_SOLVE (:110-213) */
static void gm_solve(gm_solver *s, cplx *x, const cplx *x0, const cplx *b,
                     or_gmres_info *info) {
  int M = s->M, iter = 0, jH, stat;
  int64_t n;
  double bn, sb, sPb, beta, be = -1, bea = -1;
  memset(s->H, 0, sizeof(cplx) * (M + 1) * (M + 1));
  s->iternum = 0;
  bn = gm_norm(s, b);
  if (bn < 1e-14) {
    memset(x, 0, sizeof(cplx) * s->N);
    info->stat = 1;
    info->iterations = 0;
    return;
  }
  sb = bn;
  sPb = bn;
  memcpy(x, x0, sizeof(cplx) * s->N);
  gm_apply(s, x0, s->r0);
  for (n = 0; n < s->N; n++) s->r0[n] = b[n] - s->r0[n];
  memcpy(s->w, s->r0, sizeof(cplx) * s->N);
  beta = gm_norm(s, s->w);
  if (beta / sb < s->tol) {
    info->stat = 1;
    info->be = beta / sb;
    info->bea = beta / sb;
    info->iterations = 0;
    return;
  }
  for (;;) {
    int i;
    HH(s, 0, M) = beta;
    for (n = 0; n < s->N; n++) s->V[n] = s->w[n] / beta;
    for (i = 0; i < M; i++) HH(s, i + 1, M) = 0;
    jH = 1;
    bea = gm_krylov(s, bn, &jH);
    if (bea < s->tol) {
      be = gm_final(s, jH, sPb, b, x, &stat);
      info->stat = stat;
      info->iterations = s->iternum;
      info->be = be;
      info->bea = bea;
      return;
    }
    iter++;
    if (iter >= s->maxit) break;
    beta = gm_restart(s, b, x, jH);
  }
  be = gm_final(s, jH, sPb, b, x, &stat);
  info->bea = bea;
  info->be = be;
  info->iterations = s->iternum;
  info->stat = -1;
}

void or_fgmres(int64_t N, int m_out, int m_in, int maxit, double tol,
               or_matvec_fn A, void *ctx, const double *x0, const double *b,
               double *x, or_gmres_info *info, double *bea_hist,
               int hist_len) {
  gm_solver outer, inner;
  or_gmres_info ii;
  memset(&outer, 0, sizeof(outer));
  memset(&inner, 0, sizeof(inner));
  memset(&ii, 0, sizeof(ii));
  /* INIT_SEQ_FGMRES (seq_fgmres.F90:109-141): inner MaxIt = 1, same Tol */
  gm_alloc(&inner, N, m_in, 0);
  inner.maxit = 1;
  inner.tol = tol;
  inner.A = A;
  inner.ctx = ctx;
  inner.stats = &ii;
  gm_alloc(&outer, N, m_out, 1);
  outer.maxit = maxit;
  outer.tol = tol;
  outer.inner = (struct gm_solver *)&inner;
  outer.A = A;
  outer.ctx = ctx;
  outer.stats = &ii;
  outer.hist = bea_hist;
  outer.hist_len = hist_len;
  outer.is_outer = 1;
  gm_solve(&outer, (cplx *)x, (const cplx *)x0, (const cplx *)b, &ii);
  *info = ii;
  gm_free(&outer);
  gm_free(&inner);
}

/* ------------------------------------------------------------------------ */
/* SolveEquation (ie_solver_module.F90:43-114)                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  const or_model *m;
  const double *Gspec, *csiga;
  int nthreads;
} solve_ctx;

static void solve_matvec(void *c, const double *in, double *out) {
  solve_ctx *sc = (solve_ctx *)c;
  or_apply(sc->m, sc->Gspec, sc->csiga, in, out, sc->nthreads);
}

void or_solve(const or_model *m, const double *Gspec, const double *csiga_d,
              const double *Ebkg_d, const double *E0_d, int m_out, int m_in,
              int maxit, double tol, double *Esol_d, or_gmres_info *info,
              int nthreads) {
  int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz, x, y, z, c;
  int64_t nc = (int64_t)Nx * Ny * Nz, N = 3 * nc;
  const cplx *csiga = (const cplx *)csiga_d, *Ebkg = (const cplx *)Ebkg_d,
             *E0 = (const cplx *)E0_d;
  cplx *Esol = (cplx *)Esol_d;
  cplx *rhs = malloc(sizeof(cplx) * N), *guess = malloc(sizeof(cplx) * N);
  solve_ctx sc = {m, Gspec, csiga_d, nthreads};
  for (c = 0; c < 3; c++)
    for (z = 0; z < Nz; z++) {
      cplx csigb;
      double sq;
      layer_sig(m, z, &csigb, &sq);
      for (y = 0; y < Ny; y++)
        for (x = 0; x < Nx; x++) {
          int64_t ci = x + (int64_t)Nx * (y + (int64_t)Ny * z), vi = ci + nc * c;
          rhs[vi] = sq * Ebkg[vi];
          if (E0) {
            cplx asiga = 2.0 * sq / (csiga[ci] + conj(csigb));
            guess[vi] = E0[vi] / asiga;
          } else {
            guess[vi] = rhs[vi];
          }
        }
    }
  or_fgmres(N, m_out, m_in, maxit, tol, solve_matvec, &sc, (double *)guess,
            (double *)rhs, Esol_d, info, NULL, 0);
  for (c = 0; c < 3; c++)
    for (z = 0; z < Nz; z++) {
      cplx csigb;
      double sq;
      layer_sig(m, z, &csigb, &sq);
      for (y = 0; y < Ny; y++)
        for (x = 0; x < Nx; x++) {
          int64_t ci = x + (int64_t)Nx * (y + (int64_t)Ny * z), vi = ci + nc * c;
          cplx asiga = 2.0 * sq / (csiga[ci] + conj(csigb));
          Esol[vi] = asiga * Esol[vi];
        }
    }
  free(rhs);
  free(guess);
}

/* ======================================================================== */
/* Receiver recalculation: the continuation of the IE solution to receivers */
/* (Continuation_Function_Module.F90, calc_rc_tensor_module.F90,            */
/* rc_kernel_hankel_module.F90, sources_module.F90:26-66).                  */
/* ======================================================================== */

/* PrepareRecvs (data_types_module.F90:124-152): receiver layer (from the
 * depth as given), the anomaly cell index, and depths strictly inside an
 * anomaly cell moved up to that cell's top border. */
void or_prepare_recvs(const or_model *m, int Nr, const double *depth,
                      double *zr, int *layer, int *cell) {
  int r, Iz, Nz = m->Nz;
  for (r = 0; r < Nr; r++) {
    double z = depth[r];
    int l = get_layer(m, z);
    layer[r] = l < 1 ? 1 : l;
    zr[r] = z;
    if (z <= m->z[0]) {
      cell[r] = 0;
    } else if (z >= m->z[Nz]) {
      cell[r] = Nz + 1;
    } else {
      cell[r] = -1;
      for (Iz = 1; Iz <= Nz; Iz++) {
        if (fabs(z / m->z[Iz] - 1.0) < 1e-7) {
          cell[r] = Iz;
          break;
        } else if (z > m->z[Iz - 1] && z < m->z[Iz]) {
          zr[r] = m->z[Iz - 1];
          cell[r] = Iz - 1;
          break;
        }
      }
    }
  }
}

/* PrepareExps2 (VolumeBesselTransforms.F90:342-371): the two edges of the
 * source rectangle seen from the point, in both directions. */
static void vb_prepare_exps2(double phi, double alpha, double beta, vb_edt *e) {
  int i, k;
  double cphi = cos(phi), sphi = sin(phi);
  for (i = 0; i < VB_N; i++) {
    double x[2], y[2];
    x[0] = e->explt[i] * cphi * 0.5;
    x[1] = x[0] + alpha * e->explt[i] * 0.5;
    y[0] = e->explt[i] * sphi * 0.5;
    y[1] = y[0] + beta * e->explt[i] * 0.5;
    for (k = 0; k < 2; k++) {
      e->xy[0][i][k] = x[k];
      e->xy2[0][i][k] = x[k] * x[k];
      e->xexp2[0][i][k] = exp(-(x[k] * x[k]));
      e->xerf[0][i][k] = erf(x[k]);
      e->xy[1][i][k] = y[k];
      e->xy2[1][i][k] = y[k] * y[k];
      e->xexp2[1][i][k] = exp(-(y[k] * y[k]));
      e->xerf[1][i][k] = erf(y[k]);
    }
  }
}

/* first difference f(2)-f(1) of the SINGLE_INT_* routines */
#define D1(f) ((f)[1] - (f)[0])

/* SINGLE_INT_D0 (:553-567) */
static void vb_single_int_d0(const vb_edt *e, int i, int j, double *R) {
  const double *x = e->xy[j][i], *x2 = e->xy2[j][i];
  double f[2], a, b, c;
  int k;
  for (k = 0; k < 2; k++) f[k] = e->xerf[j][i][k] * sqrt(PI);
  a = D1(f);
  R[0] = a * 0.5;
  for (k = 0; k < 2; k++) f[k] = x[k] * e->xexp2[j][i][k];
  b = D1(f);
  R[1] = a - 2.0 * b;
  for (k = 0; k < 2; k++) f[k] = f[k] * x2[k];
  c = D1(f);
  R[2] = 6.0 * a - 12.0 * b - 8.0 * c;
}

/* SINGLE_INT_D1 (:569-585) */
static void vb_single_int_d1(const vb_edt *e, int i, int j, double *R) {
  const double *x2 = e->xy2[j][i];
  double f[2];
  int k;
  for (k = 0; k < 2; k++) f[k] = e->xexp2[j][i][k];
  R[0] = D1(f) / 2.0;
  for (k = 0; k < 2; k++) f[k] = f[k] * x2[k];
  R[1] = D1(f) * 2.0;
  for (k = 0; k < 2; k++) f[k] = f[k] * x2[k];
  R[2] = D1(f) * 8.0;
}

/* SINGLE_INT_D2 (:587-603) */
static void vb_single_int_d2(const vb_edt *e, int i, int j, double *R) {
  const double *x = e->xy[j][i], *x2 = e->xy2[j][i];
  double f[2], a, b, c;
  int k;
  for (k = 0; k < 2; k++) f[k] = x[k] * e->xexp2[j][i][k];
  a = D1(f);
  R[0] = -a * 0.5;
  for (k = 0; k < 2; k++) f[k] = f[k] * x2[k];
  b = D1(f);
  R[1] = 2.0 * (a - b);
  for (k = 0; k < 2; k++) f[k] = f[k] * x2[k];
  c = D1(f);
  R[2] = 16.0 * b - 8.0 * c;
}
#undef D1

/* outfunc_int2_all (:407-434) */
static void vb_outfunc_int2(const vb_edt *e, int i, double *R) {
  double fx[3][3], fy[3][3];
  vb_single_int_d0(e, i, 0, fx[0]);
  vb_single_int_d1(e, i, 0, fx[1]);
  vb_single_int_d2(e, i, 0, fx[2]);
  vb_single_int_d0(e, i, 1, fy[0]);
  vb_single_int_d1(e, i, 1, fy[1]);
  vb_single_int_d2(e, i, 1, fy[2]);
#define CMB(a, b) (fx[a][2] * fy[b][0] + fx[a][0] * fy[b][2] + 2.0 * fx[a][1] * fy[b][1])
  R[0] = CMB(0, 0) / e->explt[i];
  R[1] = CMB(2, 0) * e->explt[i];
  R[2] = CMB(1, 1) * e->explt[i];
  R[3] = CMB(1, 0);
  R[4] = CMB(0, 2) * e->explt[i];
  R[5] = CMB(0, 1);
#undef CMB
}

/* VBTransformWeightsAllSingleInt (:198-229) + CalcWeights22 (:231-262):
 * weights of a point receiver at offset (x,y) from the source rectangle's
 * corner; WT (K,i) column-major in the order D0,DXX,DXY,DX,DYY,DY. */
void or_vbt_weights_single(double x, double y, double hx, double hy,
                           double rmin, double *WT) {
  double lms[OR_NTAPS];
  cplx f1[VB_N];
  cplx *g = (cplx *)malloc(sizeof(cplx) * VB_N * 6);
  vb_edt *e = (vb_edt *)malloc(sizeof(vb_edt));
  double rho = sqrt(x * x + y * y);
  /* scalings of :222-228: D0 times rho; DXX, DXY, DYY over rho */
  static const int rscale[6] = {+1, -1, -1, 0, -1, 0};
  int i, j;
  vb_init(rmin, lms, f1);
  vb_calc_exp_lt(log(rho) - log(rmin), e);
  vb_prepare_exps2(atan2(y, x), hx / rho, hy / rho, e);
  for (i = 0; i < VB_N; i += 2) {
    double R[6];
    vb_outfunc_int2(e, i, R);
    for (j = 0; j < 6; j++) g[j * VB_N + i] = R[j];
    vb_outfunc_int2(e, i + 1, R);
    for (j = 0; j < 6; j++) g[j * VB_N + i + 1] = -R[j];
  }
  for (j = 0; j < 6; j++) {
    cplx *h = g + j * VB_N;
    fft_c(h, VB_N, VB_FWD);
    for (i = 0; i < VB_N; i += 2) {
      h[i] = h[i] / f1[i];
      h[i + 1] = -h[i + 1] / f1[i + 1];
    }
    fft_c(h, VB_N, VB_BWD);
    for (i = 1; i < VB_N; i += 2) {
      h[i - 1] = h[i - 1] / (double)VB_N;
      h[i] = -h[i] / (double)VB_N;
    }
    for (i = OR_NFIRST; i <= OR_NLAST; i++) {
      double v = creal(h[i]);
      WT[(i - OR_NFIRST) + OR_NTAPS * j] = rscale[j] > 0 ? v * rho : (rscale[j] < 0 ? v / rho : v);
    }
  }
  free(e);
  free(g);
}

/* RC weight order (rc_kernel_hankel_module.F90:30-36) is the IE order. */
/* E entries REXX..REZZ and H entries RHXX..RHZY (const_module.F90:112-130) */
enum { REXX = 0, REXY, REXZ, REYY, REYZ, REZX, REZY, REZZ };
enum { RHXX = 0, RHXY, RHXZ, RHYX, RHYZ, RHZX, RHZY };

/* Calc_Integral_U (rc_kernel_hankel_module.F90:56-150) with Calc_Recv_Part
 * (187-230), Calc_Ftb (232-254) and Calc_G_not_match (152-186).  GE is
 * complex (8,Nz), GH complex (7,Nz), both accumulated. */
static void rc_integral_U(const or_model *m, double zrecv, int rl, int N,
                          double lms, const double *WT, cplx *GE, cplx *GH) {
  int Nz = m->Nz, Nl = m->Nl, I, l, k;
  cplx Arr[OR_MAXL + 1][2], p[OR_MAXL + 1][2], q[OR_MAXL + 1][2], eta[OR_MAXL + 1];
  cplx(*A)[OR_MAXL + 1][2] = calloc(OR_MAXL + 1, sizeof(*A));
  cplx *expz = malloc(sizeof(cplx) * (Nz + 1));
  cplx(*pez)[2] = malloc(sizeof(cplx) * 2 * (Nz + 1));
  cplx(*qez)[2] = malloc(sizeof(cplx) * 2 * (Nz + 1));
  cplx(*w1)[2] = malloc(sizeof(cplx) * 2 * (Nz + 1));
  cplx(*w2)[2] = malloc(sizeof(cplx) * 2 * (Nz + 1));
  cplx(*Ftb)[3] = calloc((size_t)3 * (Nz + 1), sizeof(cplx));
  cplx e, eta_r, eta_zr, eta_dr, iwm = m->iwm;
  double lm2 = lms * lms;
  int N1, N2, Iz;
  calc_apq(m, lms, Arr, p, q, eta, A);
  for (Iz = 1; Iz <= Nz; Iz++) expz[Iz] = cexp(-eta[m->lnum[Iz]] * (m->z[Iz] - m->z[Iz - 1]));
  /* Calc_Recv_Part */
  eta_r = eta[rl];
  eta_zr = eta_r * zrecv;
  eta_dr = rl > 1 ? eta_r * m->depth[rl - 1] * 2.0 : 0.0;
  N1 = N + 1 > 1 ? N + 1 : 1;
  N2 = N < Nz ? N : Nz;
  for (I = N1; I <= Nz; I++) {
    cplx e1, e2;
    l = m->lnum[I];
    e1 = cexp(-m->z[I - 1] * eta[l] + eta_zr);
    e2 = cexp(-m->z[I - 1] * eta[l] - eta_zr + eta_dr);
    w1[I][0] = e2 * p[rl][0] + e1;
    w1[I][1] = (-e2 * p[rl][0] + e1) * eta_r;
    w2[I][0] = e2 * p[rl][1] + e1;
    w2[I][1] = (-e2 * p[rl][1] + e1) * eta_r;
  }
  for (I = 1; I <= N2; I++) {
    cplx e1, e2;
    l = m->lnum[I];
    e1 = cexp(m->z[I] * eta[l] - eta_zr);
    /* below the last layer eta_dr is HUGE (:219): the reflected term is 0 */
    e2 = rl < Nl ? cexp(m->z[I] * eta[l] + eta_zr - eta_r * m->depth[rl] * 2.0) : 0.0;
    w1[I][0] = e2 * q[rl][0] + e1;
    w1[I][1] = (e2 * q[rl][0] - e1) * eta_r;
    w2[I][0] = e2 * q[rl][1] + e1;
    w2[I][1] = (e2 * q[rl][1] - e1) * eta_r;
  }
  /* Calc_pexpz_All / Calc_qexpz_All (apq_module.F90:112-177) */
  l = m->lnum[0];
  if (l > 1)
    e = cexp(-2.0 * eta[l] * (m->z[0] - m->depth[l - 1]));
  else if (l == 1)
    e = cexp(-2.0 * eta[1] * (m->z[0]));
  else
    e = 1.0;
  for (I = 1; I <= Nz; I++) {
    if (l != m->lnum[I]) e = 1.0;
    l = m->lnum[I];
    for (k = 0; k < 2; k++) pez[I][k] = e * p[l][k];
    e = e * expz[I] * expz[I];
  }
  l = m->lnum[Nz];
  e = (l != Nl) ? cexp(-2.0 * eta[l] * (m->depth[l] - m->z[Nz])) : 0.0;
  for (I = Nz; I >= 1; I--) {
    if (l != m->lnum[I]) e = 1.0;
    l = m->lnum[I];
    for (k = 0; k < 2; k++) qez[I][k] = e * q[l][k];
    e = e * expz[I] * expz[I];
  }
  /* Calc_Ftb: cells above the receiver use p (sign -1), below it q (+1) */
  for (I = 1; I <= Nz; I++) {
    int above = (N >= 1 && N <= Nz) ? (I <= N) : (N > Nz);
    cplx(*pq)[2] = above ? pez : qez;
    double s = above ? -1.0 : 1.0;
    cplx a, a1, e1[2], e2[2];
    l = m->lnum[I];
    a = 1.0 - expz[I];
    a1 = a / eta[l];
    for (k = 0; k < 2; k++) {
      e1[k] = 1.0 + pq[I][k] * expz[I];
      e2[k] = (1.0 - pq[I][k] * expz[I]) * s;
    }
    Ftb[I][0] = a1 * e1[0];
    Ftb[I][1] = a1 * e1[1];
    Ftb[I][2] = a * e2[1];
  }
  /* Calc_G_not_match and the accumulation (:116-145) */
  for (I = 1; I <= Nz; I++) {
    cplx a1, a2, f[3], v1[2], v2[2], sr = m->csigma[rl];
    cplx G1, G3DZ, G2DZ, G2DZ0, G2G2DZZ, G3, G1DZ, G2, G11;
    cplx *ge = GE + 8 * (I - 1), *gh = GH + 7 * (I - 1);
    l = m->lnum[I];
    a1 = A[rl][l][0];
    a2 = A[rl][l][1];
    for (k = 0; k < 3; k++) f[k] = Ftb[I][k];
    v1[0] = w1[I][0];
    v1[1] = w1[I][1];
    v2[0] = w2[I][0];
    v2[1] = w2[I][1];
    G1 = a1 * f[0] * v1[0];
    G3DZ = (a2 * f[2] * v2[1]) / sr;
    G2DZ = (a2 * f[1] * v2[1]) / sr;
    G2DZ0 = (a2 * f[2] * v2[0]) / sr;
    G2G2DZZ = lm2 * (a2 * f[1] * v2[0]) / sr;
    G3 = (a2 * f[2] * v2[0] - a1 * f[0] * v1[1]);
    G1DZ = a1 * f[0] * v1[1];
    G2 = a2 * f[1] * v2[0];
    G11 = G1 * iwm;
    ge[REXX] = ge[REXX] + G11 * WT[IE_D0] + (G3DZ + G11) * WT[IE_DXX];
    ge[REXY] = ge[REXY] + (G3DZ + G11) * WT[IE_DXY];
    ge[REXZ] = ge[REXZ] + G2DZ * WT[IE_DX];
    ge[REYY] = ge[REYY] + G11 * WT[IE_D0] + (G3DZ + G11) * WT[IE_DYY];
    ge[REYZ] = ge[REYZ] + G2DZ * WT[IE_DY];
    ge[REZX] = ge[REZX] + G2DZ0 * WT[IE_DX];
    ge[REZY] = ge[REZY] + G2DZ0 * WT[IE_DY];
    ge[REZZ] = ge[REZZ] + G2G2DZZ * WT[IE_D0];
    gh[RHXX] = gh[RHXX] + G3 * WT[IE_DXY];
    gh[RHXY] = gh[RHXY] + G3 * WT[IE_DYY] - G1DZ * WT[IE_D0];
    gh[RHXZ] = gh[RHXZ] + G2 * WT[IE_DY];
    gh[RHYX] = gh[RHYX] + G1DZ * WT[IE_D0] - G3 * WT[IE_DXX];
    gh[RHYZ] = gh[RHYZ] - G2 * WT[IE_DX];
    gh[RHZX] = gh[RHZX] - G1 * WT[IE_DY];
    gh[RHZY] = gh[RHZY] + G1 * WT[IE_DX];
  }
  free(A);
  free(expz);
  free(pez);
  free(qez);
  free(w1);
  free(w2);
  free(Ftb);
}

void or_rc_integral_U(const or_model *m, double zrecv, int layer, int cell,
                      double lm, const double *W, double *GE, double *GH) {
  rc_integral_U(m, zrecv, layer, cell, lm, W, (cplx *)GE, (cplx *)GH);
}

int64_t or_rc_block_size(int Nz, int Nr) { return 15 * (int64_t)Nr * Nz; }

/* CalcRecalculationGreenTensor3dElement (calc_rc_tensor_module.F90:164-230)
 * for one lateral offset (lx, ly) (cells, receiver minus source) and every
 * receiver.  blk: 15*Nr*Nz complex, entry e (E: 0..7, H: 8..14), receiver
 * r, source layer z at (e*Nr + r)*Nz + z. */
void or_rc_tensor_element(const or_model *m, int Nr, const double *xs,
                          const double *ys, const double *zr, const int *layer,
                          const int *cell, int lx, int ly, double *blk_d) {
  int Nz = m->Nz, r, K, c, iz;
  double rmin = sqrt(m->dx * m->dx + m->dy * m->dy) / 2.0; /* :67 */
  double lms[OR_NTAPS], WT0[OR_NTAPS * 6], WT[OR_NTAPS * 6];
  cplx *GE = malloc(sizeof(cplx) * 8 * Nz), *GH = malloc(sizeof(cplx) * 7 * Nz);
  cplx *blk = (cplx *)blk_d;
  or_vbt_lms(rmin, lms);
  for (r = 0; r < Nr; r++) {
    double x = lx * m->dx - m->dx / 2.0 + xs[r]; /* :191-192 */
    double y = ly * m->dy - m->dy / 2.0 + ys[r];
    double W0 = 0;
    or_vbt_weights_single(x, y, m->dx, m->dy, rmin, WT0);
    /* column map :196-201 (WT0 order D0,DXX,DXY,DX,DYY,DY) */
    memcpy(WT + OR_NTAPS * IE_D0, WT0 + OR_NTAPS * 0, sizeof(double) * OR_NTAPS);
    memcpy(WT + OR_NTAPS * IE_DXX, WT0 + OR_NTAPS * 1, sizeof(double) * OR_NTAPS);
    memcpy(WT + OR_NTAPS * IE_DXY, WT0 + OR_NTAPS * 2, sizeof(double) * OR_NTAPS);
    memcpy(WT + OR_NTAPS * IE_DX, WT0 + OR_NTAPS * 3, sizeof(double) * OR_NTAPS);
    memcpy(WT + OR_NTAPS * IE_DYY, WT0 + OR_NTAPS * 4, sizeof(double) * OR_NTAPS);
    memcpy(WT + OR_NTAPS * IE_DY, WT0 + OR_NTAPS * 5, sizeof(double) * OR_NTAPS);
    /* :202 W0 = MAXVAL(ABS(WT)): one scalar over all taps and columns */
    for (K = 0; K < OR_NTAPS * 6; K++)
      if (fabs(WT[K]) > W0) W0 = fabs(WT[K]);
    memset(GE, 0, sizeof(cplx) * 8 * Nz);
    memset(GH, 0, sizeof(cplx) * 7 * Nz);
    for (K = 0; K < OR_NTAPS; K++) {
      double Wm = 0, W[6], lm;
      for (c = 0; c < 6; c++) {
        double t = fabs(WT[K + OR_NTAPS * c] / W0);
        if (c == 0 || t > Wm) Wm = t;
      }
      if (Wm > WT_THRESHOLD) {
        lm = lms[K];
        W[IE_DXX] = WT[K + OR_NTAPS * IE_DXX] / lm;
        W[IE_DYY] = WT[K + OR_NTAPS * IE_DYY] / lm;
        W[IE_DXY] = WT[K + OR_NTAPS * IE_DXY] / lm;
        W[IE_DX] = WT[K + OR_NTAPS * IE_DX] * lm;
        W[IE_DY] = WT[K + OR_NTAPS * IE_DY] * lm;
        W[IE_D0] = WT[K + OR_NTAPS * IE_D0] * lm;
        rc_integral_U(m, zr[r], layer[r], cell[r], lm, W, GE, GH);
      }
    }
    /* :223-224 transpose and scale by 1/(4 pi) */
    for (iz = 0; iz < Nz; iz++) {
      for (c = 0; c < 8; c++) blk[((int64_t)c * Nr + r) * Nz + iz] = GE[c + 8 * iz] / PI / 4.0;
      for (c = 0; c < 7; c++) blk[((int64_t)(8 + c) * Nr + r) * Nz + iz] = GH[c + 7 * iz] / PI / 4.0;
    }
  }
  free(GE);
  free(GH);
}

/* CalcRecalculationGreenTensor (:37-162) on one rank: every point of the
 * 2Nx x 2Ny offset grid, p = ix*2Ny + iy with lx = ix (ix < Nx) or
 * ix - 2Nx (ix > Nx), the same for y; the ix = Nx and iy = Ny offsets are
 * zero (:70-79, :110-113). */
void or_rc_tensor_realspace(const or_model *m, int Nr, const double *xs,
                            const double *ys, const double *zr,
                            const int *layer, const int *cell, double *Greal,
                            int nthreads) {
  int Nx = m->Nx, Ny = m->Ny;
  int64_t BS = or_rc_block_size(m->Nz, Nr), p, NP = 4 * (int64_t)Nx * Ny;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (p = 0; p < NP; p++) {
    int ix = (int)(p / (2 * Ny)), iy = (int)(p % (2 * Ny));
    double *blk = Greal + 2 * BS * p;
    if (ix == Nx || iy == Ny) {
      memset(blk, 0, sizeof(cplx) * BS);
      continue;
    }
    or_rc_tensor_element(m, Nr, xs, ys, zr, layer, cell, ix < Nx ? ix : ix - 2 * Nx,
                         iy < Ny ? iy : iy - 2 * Ny, blk);
  }
}

/* CalcFFTofRCTensor (Continuation_Function_Module.F90:349-383): forward 2-D
 * DFT (exp(-i)) of every tensor column over the 2Nx x 2Ny grid. */
void or_rc_tensor_fft(const or_model *m, int Nr, const double *Greal_d,
                      double *Gspec_d, int nthreads) {
  int Nx2 = 2 * m->Nx, Ny2 = 2 * m->Ny;
  int64_t BS = or_rc_block_size(m->Nz, Nr), e;
  const cplx *Greal = (const cplx *)Greal_d;
  cplx *Gspec = (cplx *)Gspec_d;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (e = 0; e < BS; e++) {
    cplx *F = malloc(sizeof(cplx) * Nx2 * Ny2), *col = malloc(sizeof(cplx) * Nx2);
    int ix, iy;
    for (ix = 0; ix < Nx2; ix++)
      for (iy = 0; iy < Ny2; iy++) F[iy + Ny2 * ix] = Greal[e + BS * (iy + (int64_t)Ny2 * ix)];
    for (ix = 0; ix < Nx2; ix++) fft_c(F + Ny2 * ix, Ny2, -1);
    for (iy = 0; iy < Ny2; iy++) {
      for (ix = 0; ix < Nx2; ix++) col[ix] = F[iy + Ny2 * ix];
      fft_c(col, Nx2, -1);
      for (ix = 0; ix < Nx2; ix++) Gspec[e + BS * (iy + (int64_t)Ny2 * ix)] = col[ix];
    }
    free(F);
    free(col);
  }
}

/* The 17 block products of APPLY_RC_E_OP / APPLY_RC_H_OP
 * (Continuation_Function_Module.F90:408-481): (entry, input component,
 * output component 0..5 = Ex,Ey,Ez,Hx,Hy,Hz, sign).  RHYY is RHXX with
 * alpha = -1 (:468, const_module.F90:127). */
static const int RC_TERMS[17][4] = {
    {REXX, 0, 0, 1}, {REXY, 1, 0, 1}, {REXZ, 2, 0, 1},
    {REXY, 0, 1, 1}, {REYY, 1, 1, 1}, {REYZ, 2, 1, 1},
    {REZX, 0, 2, 1}, {REZY, 1, 2, 1}, {REZZ, 2, 2, 1},
    {8 + RHXX, 0, 3, 1}, {8 + RHXY, 1, 3, 1}, {8 + RHXZ, 2, 3, 1},
    {8 + RHYX, 0, 4, 1}, {8 + RHXX, 1, 4, -1}, {8 + RHYZ, 2, 4, 1},
    {8 + RHZX, 0, 5, 1}, {8 + RHZY, 1, 5, 1}};

/* LoadEint (:241-265): j = (sigma_a - sigma_b) * E, component-major. */
static cplx *rc_current(const or_model *m, const cplx *csiga, const cplx *E) {
  int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz, x, y, z, c;
  int64_t nc = (int64_t)Nx * Ny * Nz;
  cplx *J = malloc(sizeof(cplx) * 3 * nc);
  for (c = 0; c < 3; c++)
    for (z = 0; z < Nz; z++) {
      cplx csigb;
      double sq;
      layer_sig(m, z, &csigb, &sq);
      for (y = 0; y < Ny; y++)
        for (x = 0; x < Nx; x++) {
          int64_t ci = x + (int64_t)Nx * (y + (int64_t)Ny * z);
          J[ci + nc * c] = (csiga[ci] - csigb) * E[ci + nc * c];
        }
    }
  return J;
}

/* ReCalculation (:198-240): Ea, Ha complex (Nr,3,Nx,Ny) Fortran order. */
void or_rc_apply(const or_model *m, int Nr, const double *Gspec_d,
                 const double *csiga_d, const double *Eint_d, double *Ea_d,
                 double *Ha_d, int nthreads) {
  int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz, Nx2 = 2 * Nx, Ny2 = 2 * Ny;
  int64_t NS = (int64_t)Nx2 * Ny2, nc = (int64_t)Nx * Ny * Nz, s, p;
  int64_t BS = or_rc_block_size(Nz, Nr);
  const cplx *Gspec = (const cplx *)Gspec_d;
  cplx *J = rc_current(m, (const cplx *)csiga_d, (const cplx *)Eint_d);
  cplx *F = calloc((size_t)NS * 3 * Nz, sizeof(cplx)); /* [c*Nz+z][ix][iy] */
  cplx *O = calloc((size_t)NS * 6 * Nr, sizeof(cplx)); /* [c'*Nr+r][ix][iy] */
  cplx *Ea = (cplx *)Ea_d, *Ha = (cplx *)Ha_d;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (s = 0; s < 3 * Nz; s++) {
    int c = (int)(s / Nz), z = (int)(s % Nz), x, y;
    cplx *A = F + NS * s, *col = malloc(sizeof(cplx) * Nx2);
    for (x = 0; x < Nx; x++)
      for (y = 0; y < Ny; y++)
        A[y + (int64_t)Ny2 * x] = J[x + (int64_t)Nx * (y + (int64_t)Ny * z) + nc * c];
    for (x = 0; x < Nx; x++) fft_c(A + Ny2 * x, Ny2, -1);
    for (y = 0; y < Ny2; y++) {
      for (x = 0; x < Nx2; x++) col[x] = A[y + (int64_t)Ny2 * x];
      fft_c(col, Nx2, -1);
      for (x = 0; x < Nx2; x++) A[y + (int64_t)Ny2 * x] = col[x];
    }
    free(col);
  }
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (p = 0; p < NS; p++) {
    const cplx *blk = Gspec + BS * p;
    int t, r, z;
    for (t = 0; t < 17; t++) {
      int e = RC_TERMS[t][0], ci = RC_TERMS[t][1], co = RC_TERMS[t][2];
      double sg = RC_TERMS[t][3];
      for (r = 0; r < Nr; r++) {
        cplx acc = 0;
        for (z = 0; z < Nz; z++) acc += blk[((int64_t)e * Nr + r) * Nz + z] * F[p + NS * (ci * Nz + z)];
        O[p + NS * (co * Nr + r)] += sg * acc;
      }
    }
  }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (s = 0; s < 6 * Nr; s++) {
    int co = (int)(s / Nr), r = (int)(s % Nr), x, y;
    cplx *A = O + NS * s, *col = malloc(sizeof(cplx) * Nx2);
    for (x = 0; x < Nx2; x++) fft_c(A + Ny2 * x, Ny2, +1);
    for (y = 0; y < Ny; y++) {
      for (x = 0; x < Nx2; x++) col[x] = A[y + (int64_t)Ny2 * x];
      fft_c(col, Nx2, +1);
      for (x = 0; x < Nx2; x++) A[y + (int64_t)Ny2 * x] = col[x];
    }
    for (y = 0; y < Ny; y++)
      for (x = 0; x < Nx; x++) {
        cplx v = A[y + (int64_t)Ny2 * x] / (4.0 * Nx * Ny); /* RC_OP_FFTW_BWD :347 */
        int64_t o = r + (int64_t)Nr * ((co % 3) + 3 * (x + (int64_t)Nx * y));
        if (co < 3)
          Ea[o] = v;
        else
          Ha[o] = v;
      }
    free(col);
  }
  free(J);
  free(F);
  free(O);
}

/* Direct sum of the convolution the FFT path computes: no FFT, no padding. */
void or_rc_apply_bruteforce(const or_model *m, int Nr, const double *Greal_d,
                            const double *csiga_d, const double *Eint_d,
                            double *Ea_d, double *Ha_d, int nthreads) {
  int Nx = m->Nx, Ny = m->Ny, Nz = m->Nz;
  int64_t nc = (int64_t)Nx * Ny * Nz, BS = or_rc_block_size(Nz, Nr), q;
  const cplx *Greal = (const cplx *)Greal_d;
  cplx *J = rc_current(m, (const cplx *)csiga_d, (const cplx *)Eint_d);
  cplx *Ea = (cplx *)Ea_d, *Ha = (cplx *)Ha_d;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (q = 0; q < (int64_t)Nx * Ny; q++) {
    int xr = (int)(q % Nx), yr = (int)(q / Nx), xs, ys, t, r, z;
    cplx acc[6 * 64];
    for (t = 0; t < 6 * Nr; t++) acc[t] = 0;
    for (ys = 0; ys < Ny; ys++)
      for (xs = 0; xs < Nx; xs++) {
        int lx = xr - xs, ly = yr - ys;
        int ix = lx >= 0 ? lx : lx + 2 * Nx, iy = ly >= 0 ? ly : ly + 2 * Ny;
        const cplx *blk = Greal + BS * (iy + (int64_t)2 * Ny * ix);
        for (t = 0; t < 17; t++) {
          int e = RC_TERMS[t][0], ci = RC_TERMS[t][1], co = RC_TERMS[t][2];
          double sg = RC_TERMS[t][3];
          for (r = 0; r < Nr; r++)
            for (z = 0; z < Nz; z++)
              acc[co * Nr + r] += sg * blk[((int64_t)e * Nr + r) * Nz + z] *
                                  J[xs + (int64_t)Nx * (ys + (int64_t)Ny * z) + nc * ci];
        }
      }
    for (t = 0; t < 6 * Nr; t++) {
      int co = t / Nr, r = t % Nr;
      int64_t o = r + (int64_t)Nr * ((co % 3) + 3 * (xr + (int64_t)Nx * yr));
      if (co < 3)
        Ea[o] = acc[t];
      else
        Ha[o] = acc[t];
    }
  }
  free(J);
}

/* PlaneWave (sources_module.F90:26-66): the normal field at receiver depths
 * for one polarization (0 = EX, 1 = EY).  F0 complex (Nr,6), components
 * Ex,Ey,Ez,Hx,Hy,Hz. */
void or_plane_wave(const or_model *m, int pol, int Nr, const double *zrecv,
                   double *F0_d) {
  int Nl = m->Nl, r, l;
  cplx Arr[OR_MAXL + 1][2], p[OR_MAXL + 1][2], q[OR_MAXL + 1][2], eta[OR_MAXL + 1], t1;
  cplx(*A)[OR_MAXL + 1][2] = calloc(OR_MAXL + 1, sizeof(*A));
  cplx *F0 = (cplx *)F0_d;
  memset(F0, 0, sizeof(cplx) * 6 * Nr);
  calc_apq(m, 0.0, Arr, p, q, eta, A);
  t1 = Nl > 1 ? 1.0 + q[1][0] * cexp(-m->depth[1] * eta[1] * 2.0) : 1.0;
  t1 = A[1][1][0] * t1;
  for (r = 0; r < Nr; r++) {
    cplx t2, e1, e2, G0, G1;
    l = get_layer(m, zrecv[r]);
    l = l < 1 ? 1 : l;
    t2 = A[l][1][0] / t1;
    e1 = cexp(-eta[l] * zrecv[r]);
    e2 = l < Nl ? q[l][0] * cexp(-eta[l] * (2.0 * m->depth[l] - zrecv[r])) : 0.0;
    G0 = t2 * (e1 + e2);
    G1 = -t2 * (e1 - e2) * eta[l] / m->iwm;
    if (pol == 0) {
      F0[r + Nr * 0] = G0;
      F0[r + Nr * 4] = G1;
    } else {
      F0[r + Nr * 1] = G0;
      F0[r + Nr * 3] = -G1;
    }
  }
  free(A);
}
