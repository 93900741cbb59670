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
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "g2g_common.hpp"
#include "g2g_fft.hpp"

namespace g2g {

// Operator counters (the reference's split: Integral_Equation_Module.F90:
// 44-56, ie_solver_module.F90:98-113).  Times of the FFT passes, the
// contraction and the Krylov dot products come from HIP events and are
// gathered while profiling is on (giem2g_op_set_profiling) or from
// giem2g_op_time_apply; counts are always kept.
struct Counters {
  int64_t matvecs = 0, matvecs32 = 0, dotprods = 0;
  double t_apply = 0, t_fwd = 0, t_contr = 0, t_bwd = 0, t_dot = 0;
  double t_tensor = 0, t_tensor_fft = 0;
  // per precision (0 FP64, 1 complex64): contraction launches and their
  // total time, and how many right-hand sides those launches served
  int64_t contr_launches[2] = {0, 0}, contr_rhs[2] = {0, 0};
  double t_contr_p[2] = {0, 0};
  // classical Gram-Schmidt passes and the bytes the Krylov kernels move
  // (multi-dot, update, norms, linear combinations, copies, conversions)
  int64_t gs_passes = 0;
  double krylov_bytes = 0;
};

// Event log of the profiling mode: pairs of HIP events recorded around the
// stages of each apply and around each Gram-Schmidt pass, read back (one
// synchronisation) by flush().  Shared by the threads of a paired solve.
struct Prof {
  enum Kind { FWD = 0, CONTR = 1, BWD = 2, DOT = 3 };
  bool on = false;
  std::mutex mu;
  std::vector<hipEvent_t> pool;
  struct Mark {
    int kind, prec, nrhs;
    int a, b;
  };
  std::vector<Mark> marks;
  size_t next = 0;
  // record one event on stream s, returning its index in the pool
  int record(hipStream_t s) {
    std::lock_guard<std::mutex> lk(mu);
    if (next == pool.size()) {
      hipEvent_t e;
      G2G_HIP(hipEventCreate(&e));
      pool.push_back(e);
    }
    G2G_HIP(hipEventRecord(pool[next], s));
    return (int)next++;
  }
  void mark(int kind, int prec, int nrhs, int a, int b) {
    std::lock_guard<std::mutex> lk(mu);
    marks.push_back(Mark{kind, prec, nrhs, a, b});
  }
  // wait for the recorded events and add their intervals to c; called
  // between API calls only (never while a solve still records)
  void flush(Counters& c) {
    std::lock_guard<std::mutex> lk(mu);
    for (const Mark& m : marks) {
      G2G_HIP(hipEventSynchronize(pool[m.b]));
      float ms = 0;
      G2G_HIP(hipEventElapsedTime(&ms, pool[m.a], pool[m.b]));
      const double t = ms * 1e-3;
      switch (m.kind) {
        case FWD: c.t_fwd += t; c.t_apply += t; break;
        case CONTR:
          c.t_contr += t;
          c.t_apply += t;
          c.t_contr_p[m.prec] += t;
          c.contr_launches[m.prec]++;
          c.contr_rhs[m.prec] += m.nrhs;
          break;
        case BWD: c.t_bwd += t; c.t_apply += t; break;
        default: c.t_dot += t; break;
      }
    }
    marks.clear();
    next = 0;
  }
  void release() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
    marks.clear();
    next = 0;
  }
};

// Device buffers that repeated solves reuse: each slot keeps the largest
// allocation asked of it until release(), so a solve after the first of
// the same size allocates and frees nothing (at 256x256x32 the Krylov
// bases of one solve are several GB).
struct Arena {
  struct Buf {
    void* p = nullptr;
    size_t n = 0;
  };
  std::vector<Buf> b;
  void* get(int slot, size_t bytes) {
    if ((size_t)slot >= b.size()) b.resize(slot + 1);
    Buf& x = b[slot];
    if (x.n < bytes) {
      if (x.p) G2G_HIP(hipFree(x.p));
      x.p = nullptr;
      x.n = 0;
      G2G_HIP(hipMalloc(&x.p, bytes));
      x.n = bytes;
    }
    return x.p;
  }
  // pinned host buffers (coherent, mapped: kernels write them directly and
  // the host reads them without a copy), zeroed when allocated
  std::vector<Buf> h;
  void* get_pinned(int slot, size_t bytes) {
    if ((size_t)slot >= h.size()) h.resize(slot + 1);
    Buf& x = h[slot];
    if (x.n < bytes) {
      if (x.p) G2G_HIP(hipHostFree(x.p));
      x.p = nullptr;
      x.n = 0;
      G2G_HIP(hipHostMalloc(&x.p, bytes, hipHostMallocCoherent | hipHostMallocMapped));
      std::memset(x.p, 0, bytes);
      x.n = bytes;
    }
    return x.p;
  }
  size_t bytes() const {
    size_t t = 0;
    for (const Buf& x : b) t += x.n;
    return t;
  }
  void release() {
    for (Buf& x : b)
      if (x.p) (void)hipFree(x.p);
    b.clear();
    for (Buf& x : h)
      if (x.p) (void)hipHostFree(x.p);
    h.clear();
  }
};

template <typename R> struct OpBuffers {
  using T2 = cx_t<R>;
  T2* G = nullptr;      // spectral tensor, Nx*Ny*BS
  T2* X = nullptr;      // nchunk*Ny*2Nx*NR
  T2* S = nullptr;      // 2Nx*2Ny*3Nz
  T2* SB = nullptr;     // spectra of batched applies, sb_cap slots (allocated on first use)
  int sb_cap = 0;
  T2* gsig = nullptr;   // Nx*Ny*Nz
  T2* asiga = nullptr;  // Nx*Ny*Nz
  R* sq = nullptr;      // Nz: sqrt(Re sigma_b)
  R* dzn = nullptr;     // Nz: 1/(dz * 4 Nx Ny)
  T2* twx = nullptr;    // 2Nx twiddles
  T2* twy = nullptr;    // 2Ny twiddles
  int NR = 0, nchunk = 0;
  bool half = false;    // half-length FFT passes
  size_t bytes = 0;
};

struct Operator {
  int Nx = 0, Ny = 0, Nz = 0;
  double dx = 0, dy = 0;
  std::vector<double> z, dz;
  int comp_slot[3] = {0, 1, 2};  // memory slot of physical component x,y,z
  FftPlan px, py;                // factor plans of the 2Nx- and 2Ny-point transforms
  // background of the current frequency
  int Nl = 0;
  double omega = 0;
  std::vector<zd> csigma;  // 0..Nl
  std::vector<double> thick;    // 1..Nl-1 (index 0 unused)
  std::vector<int> lnum;        // 0..Nz
  std::vector<zd> csigb;   // per anomaly layer
  std::vector<double> sqsigb;
  bool have_bkg = false, have_tensor = false, have_sigma = false, have_g32 = false;
  int bkg_version = 0;  // incremented by every background (frequency) change
  // the anomaly conductivity last set (complex (Nx,Ny,Nz)): the per-cell
  // scalings derive from it and the background, so a background change
  // recomputes them from this copy, as the reference derives them from the
  // current csigb at every apply (apply_ie_operator_module.F90:66-68)
  zd* csiga = nullptr;
  Prof prof;
  int64_t BS = 0;
  OpBuffers<double> d64;
  OpBuffers<float> d32;
  hipStream_t stream = nullptr;
  int contraction = 0;  // 0 automatic, 1 vector ALU two-deep, 2 matrix cores, 3 vector ALU one-deep, 4 small-block matrix cores, 5 the same with the x-y mirror
  bool tensor_own = false;  // the spectral tensor was built here (not put by the caller)
  Counters cnt;
  // tensor-build scratch (frequency independent weights are cached)
  double* wts = nullptr;  // 6 * 501 * Nx*Ny masked, lambda-scaled weights
  bool have_wts = false;
  double* tabs = nullptr;  // per-lambda kernel tables
  zd* adj = nullptr;       // adjoint-apply scratch (allocated on first use)
  // Krylov workspaces of the solves: one per concurrent solve (a paired
  // solve runs two); kept between solves, freed with the operator
  Arena ws[2];
  size_t tabs_bytes = 0;
  hipEvent_t ev[8] = {};

  int64_t ncell() const { return (int64_t)Nx * Ny * Nz; }
  int64_t nvec() const { return 3 * ncell(); }
};

// Continuation of the IE solution to receivers (reference: RC_OPERATOR,
// Continuation_Function_Module.F90:46-80, single rank).  It shares the IE
// operator's grid, background and stream.  Tensor layout: per point
// p = ix*2Ny + iy of the full 2Nx x 2Ny grid (real-space offsets or spectrum)
// a block of BS = 15*Nr*Nz complex, entry e (0..7 = REXX, REXY, REXZ, REYY,
// REYZ, REZX, REZY, REZZ; 8..14 = RHXX, RHXY, RHXZ, RHYX, RHYZ, RHZX, RHZY),
// receiver r, source layer z at (e*Nr + r)*Nz + z.
struct RcOperator {
  Operator* op = nullptr;
  int Nr = 0;
  std::vector<double> xs, ys, depth;  // as given (lateral shifts from the cell centres)
  std::vector<double> zr;             // after PrepareRecvs
  std::vector<int> layer, cell;
  int64_t BS = 0, NP = 0;
  zd* G = nullptr;        // tensor, NP*BS
  double* wts = nullptr;  // masked, lambda-scaled weights [r][6][501][NP] (frequency independent)
  zd* tabs = nullptr;     // per-tap kernel tables of one receiver [501][13][Nz]
  zd* dsig = nullptr;     // sigma_a - sigma_b per cell
  zd *Xf = nullptr, *S = nullptr, *So = nullptr, *Xi = nullptr;  // FFT work buffers
  int NR = 0, nchunk_f = 0, nchunk_i = 0;
  bool have_wts = false, have_tensor = false, spectral = false, have_sigma = false;
  int tensor_version = -1, sigma_version = -1;  // background version each was made for
  double t_tensor = 0, t_fft = 0, t_apply = 0;
  int64_t applies = 0;
  size_t bytes = 0;
};

// GetLayerSingle (data_types_module.F90:75-103): background layer of depth z
// (0 at the surface, -1 in the air) for the operator's current background.
inline int op_layer(const Operator& op, double z) {
  if (std::fabs(z) < 1e-15) return 0;
  if (z < 0) return -1;
  if (op.Nl == 1) return 1;
  std::vector<double> depth(op.Nl + 1, 0.0);
  for (int i = 1; i <= op.Nl - 1; i++) depth[i] = depth[i - 1] + op.thick[i];
  if (z > depth[op.Nl - 1]) return op.Nl;
  int l = 1;
  while (z > depth[l]) l++;
  return l;
}

// apply kernels (g2g_apply.hip)
template <typename R> void op_alloc_buffers(Operator& op, OpBuffers<R>& b, bool with_tensor);
template <typename R> void op_free_buffers(OpBuffers<R>& b);
// mirror: the one-vector contraction may take the x-y mirror where it
// applies (op_mirror_ok); the batched applies pass false so that their
// products do not depend on how the vectors are grouped
template <typename R>
void op_apply(Operator& op, OpBuffers<R>& b, const cx_t<R>* vin, cx_t<R>* vout, hipStream_t s,
              bool timed, bool mirror = true);
// two vectors through one tensor read; needs op_pair_supported(op)
template <typename R>
void op_apply2(Operator& op, OpBuffers<R>& b, const cx_t<R>* vin0, const cx_t<R>* vin1, cx_t<R>* vout0,
               cx_t<R>* vout1, hipStream_t s);
// Nz for which the matrix-core contraction exists (8 to 64; padded to a multiple of 16)
bool op_pair_supported(const Operator& op);
// The x-y mirror of the one-vector contraction applies: a square lateral grid
// with dx = dy and a tensor the library built (a caller's tensor may lack the
// symmetry that swapping the x and y axes gives a layered earth)
inline bool op_mirror_ok(const Operator& op) { return op.tensor_own && op.Nx == op.Ny && op.dx == op.dy; }
// right-hand sides one tensor read serves: 8 (Nz 8 to 32), 2 (Nz 33 to 64), else 1
int op_batch_max(const Operator& op);
// nrhs vectors stored one after another, batches of op_batch_max (FP64: at
// most 4) per tensor read
template <typename R>
void op_apply_batch(Operator& op, OpBuffers<R>& b, int nrhs, const cx_t<R>* vin, cx_t<R>* vout,
                    hipStream_t s);
// Allocation-free batched apply: the same products as op_apply_batch with
// the FFT intermediate and the spectrum slots in a file-scope device array
// of NOALLOC_SCRATCH_BYTES per GPU (batches of as many vectors per tensor
// read as fit in it).  op_noalloc_setup(device) looks up that device's copy;
// op_init calls it with the operator's device current.  Calls that overlap
// on one GPU share the array and race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
void op_noalloc_setup(int device);
// vectors per tensor read the array allows (0 when one vector does not fit;
// *need1 = bytes one vector needs)
template <typename R> int op_noalloc_batch_max(const Operator& op, const OpBuffers<R>& b, size_t* need1);
template <typename R>
void op_apply_batch_noalloc(Operator& op, OpBuffers<R>& b, int nrhs, const cx_t<R>* vin, cx_t<R>* vout,
                            hipStream_t s);
// per-cell scalings from the conductivity d_csiga (HBM, complex (Nx,Ny,Nz));
// keeps a copy in op.csiga (d_csiga may be op.csiga itself)
void op_set_sigma(Operator& op, const zd* d_csiga);
// new cell sizes and layer borders (same Nx, Ny, Nz): the 1/(dz 4NxNy)
// factors of the inverse transform are uploaded again and the cached
// filter weights (which depend on dx, dy) are dropped
void op_set_geometry(Operator& op, double dx, double dy, const double* z);
// out = A^H in (FP64), from the reciprocity of the contraction kernel
void op_apply_adjoint(Operator& op, const zd* vin, zd* vout, hipStream_t s);
void op_make_g32(Operator& op);
// tensor build (g2g_tensor.hip)
void op_build_tensor(Operator& op);
void op_tensor_fft(Operator& op, zd* G);
void op_free_tensor_scratch(Operator& op);
// receiver continuation
void rc_init(RcOperator& rc, Operator& op, int Nr, const double* xs, const double* ys,
             const double* depth);                                     // g2g_apply.hip
void rc_fini(RcOperator& rc);                                          // g2g_apply.hip
void rc_prepare_receivers(RcOperator& rc);                             // g2g_tensor.hip
void rc_build_tensor(RcOperator& rc);                                  // g2g_tensor.hip
void rc_tensor_fft(RcOperator& rc);                                    // g2g_tensor.hip
void rc_set_sigma(RcOperator& rc, const zd* d_csiga);                  // g2g_apply.hip
void rc_apply(RcOperator& rc, const zd* E, zd* Ea, zd* Ha, hipStream_t s);  // g2g_apply.hip

}  // namespace g2g
