// g2g_op.hpp -- the device-resident IE operator object.
//
// One object holds everything the matvec A*v of the contracting integral
// equation needs on one GPU (reference: IntegralEquationOperator,
// Integral_Equation_Module.F90:59-103, single rank, no MPI split):
//   - the spectral Green tensor for kx in [0,Nx), ky in [0,Ny) in blocks of
//     BS = 4*Nz*(Nz+1)/2 + 2*Nz*Nz complex (packed-symmetric xx,xy,yy,zz then
//     column-major xz,yz), block b = kx + Nx*ky;
//   - per-cell scalings gsig = (sa-sb)*a and asiga = a, a = 2 sqrt(Re sb)/(sa+conj sb)
//     (apply_ie_operator_module.F90:66-68, 134);
//   - FFT work buffers: X (x-transformed rows, tiled by NR row chunks) and
//     S (full 2Nx x 2Ny spectrum, [kx][ky][c*Nz+z]).
#pragma once
#include <vector>

#include "g2g_common.hpp"

namespace g2g {

struct Counters {
  int64_t matvecs = 0, matvecs32 = 0, dotprods = 0;
  double t_apply = 0, t_fwd = 0, t_contr = 0, t_bwd = 0, t_dot = 0;
  double t_tensor = 0, t_tensor_fft = 0;
};

template <typename R> struct OpBuffers {
  using T2 = cx_t<R>;
  T2* G = nullptr;      // spectral tensor, Nx*Ny*BS
  T2* X = nullptr;      // nchunk*Ny*2Nx*NR
  T2* S = nullptr;      // 2Nx*2Ny*3Nz
  T2* S2 = nullptr;     // second spectrum for two-vector applies (allocated on first use)
  T2* gsig = nullptr;   // Nx*Ny*Nz
  T2* asiga = nullptr;  // Nx*Ny*Nz
  R* sq = nullptr;      // Nz: sqrt(Re sigma_b)
  R* dzn = nullptr;     // Nz: 1/(dz * 4 Nx Ny)
  T2* twx = nullptr;    // 2Nx twiddles
  T2* twy = nullptr;    // 2Ny twiddles
  int NR = 0, nchunk = 0;
  size_t bytes = 0;
};

struct Operator {
  int Nx = 0, Ny = 0, Nz = 0;
  double dx = 0, dy = 0;
  std::vector<double> z, dz;
  int comp_slot[3] = {0, 1, 2};  // memory slot of physical component x,y,z
  // background of the current frequency
  int Nl = 0;
  double omega = 0;
  std::vector<zd> csigma;  // 0..Nl
  std::vector<double> thick;    // 1..Nl-1 (index 0 unused)
  std::vector<int> lnum;        // 0..Nz
  std::vector<zd> csigb;   // per anomaly layer
  std::vector<double> sqsigb;
  bool have_bkg = false, have_tensor = false, have_sigma = false, have_g32 = false;
  int64_t BS = 0;
  OpBuffers<double> d64;
  OpBuffers<float> d32;
  hipStream_t stream = nullptr;
  int contraction = 0;  // 0 automatic (= 1), 1 vector ALU two-deep, 2 matrix cores, 3 vector ALU one-deep
  Counters cnt;
  // tensor-build scratch (frequency independent weights are cached)
  double* wts = nullptr;  // 6 * 501 * Nx*Ny masked, lambda-scaled weights
  bool have_wts = false;
  double* tabs = nullptr;  // per-lambda kernel tables
  zd* adj = nullptr;       // adjoint-apply scratch (allocated on first use)
  size_t tabs_bytes = 0;
  hipEvent_t ev[8] = {};

  int64_t ncell() const { return (int64_t)Nx * Ny * Nz; }
  int64_t nvec() const { return 3 * ncell(); }
};

// apply kernels (g2g_apply.hip)
template <typename R> void op_alloc_buffers(Operator& op, OpBuffers<R>& b, bool with_tensor);
template <typename R> void op_free_buffers(OpBuffers<R>& b);
template <typename R>
void op_apply(Operator& op, OpBuffers<R>& b, const cx_t<R>* vin, cx_t<R>* vout, hipStream_t s,
              bool timed);
template <typename R>
void op_apply2(Operator& op, OpBuffers<R>& b, const cx_t<R>* vin0, const cx_t<R>* vin1, cx_t<R>* vout0,
               cx_t<R>* vout1, hipStream_t s);
void op_set_sigma(Operator& op, const zd* d_csiga);
// out = A^H in (FP64), from the reciprocity of the contraction kernel
void op_apply_adjoint(Operator& op, const zd* vin, zd* vout, hipStream_t s);
void op_make_g32(Operator& op);
// tensor build (g2g_tensor.hip)
void op_build_tensor(Operator& op);
void op_tensor_fft(Operator& op, zd* G);
void op_free_tensor_scratch(Operator& op);

}  // namespace g2g
