// g2g_apply.hip -- the IE operator matvec on one MI355X.
//
// Reference path replaced: APPLY_PRECOND -> MULT_IE_OP -> {distributed FFT,
// VERTICAL_MULT_GENERAL(_ZERO), inverse FFT} (apply_ie_operator_module.F90:
// 39-374, Distributed_FFT.F90:220-238,437-505,639-788).  Five kernels:
//   K1 fwd_x   : pre-scale by gsig, zero-pad x to 2Nx, FFT along x (LDS),
//                write X tiled [chunk][y][kx][row]
//   K2 fwd_y   : per kx column, zero-pad y to 2Ny, FFT along y, write the
//                spectrum S[kx][ky][c*Nz+z] (Nyquist row/column skipped)
//   K3 contract: per (kx,ky) quadrant block, the 3Nz x 3Nz block product for
//                the four points (+-kx,+-ky) sharing one stored tensor block,
//                staged through LDS; Nyquist points written as zero
//   K4 inv_y   : inverse FFT along y, keep y < Ny
//   K5 inv_x   : inverse FFT along x, keep x < Nx, post-scale
//                out = sqrt(Re sb) * (asiga*v - F/(dz*4NxNy))
// Each kernel is templated on the real type (double: the FP64 drop-in;
// float: the complex64 inner-solver operator).
#include "g2g_fft.hpp"
#include "g2g_op.hpp"

#include <cmath>
#include <type_traits>
#include <utility>

namespace g2g {

constexpr int NT = 256;  // threads per block for every apply kernel

// Rows per FFT block: at most 32 KB of LDS per block (four blocks per CU)
// and at most 8 rows.  Measured: 64 KB blocks (two per CU) were 17-30%
// slower; at 256-point complex64 rows 8 rows per block were 9-26% faster
// than 16; 4 rows (32 B contiguous per transposed row segment in complex64,
// 64 B in FP64) were 9-29% slower than 8.
template <typename R> constexpr int nr_for(int L) {
  constexpr int kBytes = 32 * 1024;
  const int n = kBytes / (L * 2 * (int)sizeof(R));
  return n < 1 ? 1 : (n > 8 ? 8 : n);
}
// rows per block of a pass over rows of length L: that of L's buffer
// capacity (L itself for powers of two, the next power of two otherwise)
template <typename R> int nr_rt(int L) { return nr_for<R>(len_capacity(L)); }
template <typename R> int nr_half_rt(int L) { return nr_for<R>(L / 2); }
// Nz for which the matrix-core contraction of nrhs right-hand sides exists:
// up to 64 layers for one or two, up to 32 for four or eight; at least C3_MIN_NZ
// (a layer count that is not a multiple of 16 is padded to the next one)
constexpr int C3_MIN_NZ = 8;
inline bool c3_supported(int nz, int nrhs) { return nz >= C3_MIN_NZ && nz <= (nrhs <= 2 ? 64 : 32); }
// Half-length passes (two N-point transforms per zero-padded 2N-point one)
// for rows of 32 points or more.  Measured against the 2N-point passes: at
// 256x256 (512-point rows) the FP64 and complex64 passes together 15% faster,
// at 128x128 FP64 5-7% faster and complex64 within 4%.  Powers of two only.
template <typename R> bool use_half_fft(int Lx, int Ly) {
  return Lx >= 32 && Ly >= 32 && is_pow2(Lx) && is_pow2(Ly);
}

struct Geo {
  int Nx, Ny, Nz, Nz3;
  int nrx_log;  // X tile rows = 1 << nrx_log
  int64_t nc;   // cells
  int slot[3];
  FftPlan px, py;  // factors of 2Nx and 2Ny (used where they are not powers of two)
  // x passes: chunks of NR rows per field component when NR divides Nz
  // (then the three components of a layer range run as consecutive blocks
  // and share the per-cell scalings in cache), else 0
  int chz;
};

// x-pass chunk of block row by: with chz > 0 the blocks run component-
// interleaved (by = 3 zc + c), so the reads of gsig / asiga for one layer
// range by the three components follow each other
__device__ inline int x_chunk(const Geo& g, int by) {
  if (g.chz == 0) return by;
  const int zc = by / 3, c = by - 3 * zc;
  return c * g.chz + zc;
}

// y passes: column kx and row chunk of block (blockIdx.x, blockIdx.y).
// Workgroups go round-robin to the eight XCDs, each with its own L2; here
// the four workgroups that one XCD runs one after another take a 2x2 group
// of (kx pair, chunk pair).  Neighbouring columns share 128-B lines of X,
// and neighbouring chunks share lines of S, when a chunk's rows fill only
// 64 B (eight complex64 rows): the two halves of a line then meet in one
// L2.  A bijection when gridDim.x and gridDim.y are even and their product
// is a multiple of 32; otherwise, or when PAIR is false (rows that fill whole
// lines: FP64, where the remap measured 2-8% slower), the plain mapping.
template <bool PAIR> __device__ inline void y_tile(int& kx, int& chunk) {
  const int B = gridDim.x, C = gridDim.y;
  if (!PAIR || (B & 1) || (C & 1) || (B * C) % 32) {
    kx = blockIdx.x;
    chunk = blockIdx.y;
    return;
  }
  const int b = blockIdx.y * B + blockIdx.x;
  const int xcd = b & 7, t = b >> 3;
  const int st = (t >> 2) * 8 + xcd, r = t & 3;
  const int hb = B >> 1;
  kx = 2 * (st % hb) + (r & 1);
  chunk = 2 * (st / hb) + (r >> 1);
}

__device__ inline int64_t xidx(const Geo& g, int cz, int y, int kx) {
  const int nrx = 1 << g.nrx_log;
  return ((((int64_t)(cz >> g.nrx_log) * g.Ny + y) * (2 * g.Nx) + kx) << g.nrx_log) + (cz & (nrx - 1));
}

// Tensor elements are loaded with ld_stream (each apply reads them once):
// FP64, 128x128x32, contraction 6% and inverse y pass 17% faster.

template <typename T2>
__device__ inline void load_tw(T2* tw, const T2* g, int L) {
  for (int i = threadIdx.x; i < L; i += NT) tw[i] = g[i];
}

// ---------------------------------------------------------------- K1 ----
// L is the buffer capacity; with RT the rows are 2Nx <= L points long
// (a length that is not a power of two, transformed by its factor plan).
template <typename R, int L, bool RT = false>
__global__ __launch_bounds__(NT) void k_fwd_x(const cx_t<R>* __restrict__ v,
                                              const cx_t<R>* __restrict__ gsig,
                                              cx_t<R>* __restrict__ X,
                                              const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int NR = nr_for<R>(L);
  constexpr int LS = L + 1;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 tw[L];
  const int y = blockIdx.x, chunk = blockIdx.y, Nx = g.Nx;
  const int Lr = RT ? 2 * Nx : L;
  load_tw<T2>(tw, twg, Lr);
  for (int i = threadIdx.x; i < NR * Nx; i += NT) {
    const int s = i / Nx, x = i - s * Nx, cz = chunk * NR + s;
    T2 val = mk<T2>(0, 0);
    if (cz < g.Nz3) {
      const int c = cz / g.Nz, z = cz - c * g.Nz;
      const int64_t ci = ((int64_t)z * g.Ny + y) * Nx + x;
      val = v[g.slot[c] * g.nc + ci] * gsig[ci];
    }
    buf[s * LS + x] = val;
    buf[s * LS + x + Nx] = mk<T2>(0, 0);
  }
  __syncthreads();
  lds_fft_any<T2, L, NR, LS, NT, -1, RT>(buf, tw, g.px);
  // X tile for this chunk is contiguous: [y][kx][s]
  T2* dst = X + (((int64_t)chunk * g.Ny + y) * Lr) * NR;
  for (int i = threadIdx.x; i < NR * Lr; i += NT) {
    const int kx = i / NR, s = i - kx * NR;
    dst[i] = buf[s * LS + kx];
  }
}

// ---------------------------------------------------------------- K2 ----
// FULL: every point of the 2Nx x 2Ny spectrum (the receiver continuation);
// otherwise the x- and y-Nyquist lines, which the IE operator never uses,
// are skipped.
template <typename R, int L, bool FULL = false, bool RT = false>
__global__ __launch_bounds__(NT) void k_fwd_y(const cx_t<R>* __restrict__ X,
                                              cx_t<R>* __restrict__ S,
                                              const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int NR = nr_for<R>(L);
  constexpr int LS = L + 1;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 tw[L];
  const int kx = blockIdx.x, chunk = blockIdx.y, Ny = g.Ny;
  const int Lr = RT ? 2 * Ny : L;
  if (!FULL && kx == g.Nx) return;  // x-Nyquist column is never used
  load_tw<T2>(tw, twg, Lr);
  for (int i = threadIdx.x; i < NR * Ny; i += NT) {
    const int yy = i / NR, s = i - yy * NR, cz = chunk * NR + s;
    buf[s * LS + yy] = (cz < g.Nz3) ? X[xidx(g, cz, yy, kx)] : mk<T2>(0, 0);
    buf[s * LS + yy + Ny] = mk<T2>(0, 0);
  }
  __syncthreads();
  lds_fft_any<T2, L, NR, LS, NT, -1, RT>(buf, tw, g.py);
  for (int i = threadIdx.x; i < NR * Lr; i += NT) {
    const int ky = i / NR, s = i - ky * NR, cz = chunk * NR + s;
    if ((FULL || ky != Ny) && cz < g.Nz3) S[((int64_t)kx * Lr + ky) * g.Nz3 + cz] = buf[s * LS + ky];
  }
}

// ---------------------------------------------------------------- K4 ----
template <typename R, int L, bool FULL = false, bool RT = false>
__global__ __launch_bounds__(NT) void k_inv_y(const cx_t<R>* __restrict__ S,
                                              cx_t<R>* __restrict__ X,
                                              const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int NR = nr_for<R>(L);
  constexpr int LS = L + 1;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 tw[L];
  const int kx = blockIdx.x, chunk = blockIdx.y, Ny = g.Ny;
  const int Lr = RT ? 2 * Ny : L;
  if (!FULL && kx == g.Nx) {  // x-Nyquist: zero (InitialTransposeRepack, Distributed_FFT.F90:723-724)
    for (int i = threadIdx.x; i < NR * Ny; i += NT) {
      const int yy = i / NR, s = i - yy * NR, cz = chunk * NR + s;
      if (cz < g.Nz3) X[xidx(g, cz, yy, kx)] = mk<T2>(0, 0);
    }
    return;
  }
  load_tw<T2>(tw, twg, Lr);
  for (int i = threadIdx.x; i < NR * Lr; i += NT) {
    const int ky = i / NR, s = i - ky * NR, cz = chunk * NR + s;
    buf[s * LS + ky] = (cz < g.Nz3) ? S[((int64_t)kx * Lr + ky) * g.Nz3 + cz] : mk<T2>(0, 0);
  }
  __syncthreads();
  lds_fft_any<T2, L, NR, LS, NT, +1, RT>(buf, tw, g.py);
  for (int i = threadIdx.x; i < NR * Ny; i += NT) {
    const int yy = i / NR, s = i - yy * NR, cz = chunk * NR + s;
    if (cz < g.Nz3) X[xidx(g, cz, yy, kx)] = buf[s * LS + yy];
  }
}

// ---------------------------------------------------------------- K5 ----
template <typename R, int L, bool RT = false>
__global__ __launch_bounds__(NT) void k_inv_x(const cx_t<R>* __restrict__ X,
                                              const cx_t<R>* __restrict__ vin,
                                              const cx_t<R>* __restrict__ asiga,
                                              const R* __restrict__ sq,
                                              const R* __restrict__ dzn,
                                              cx_t<R>* __restrict__ vout,
                                              const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int NR = nr_for<R>(L);
  constexpr int LS = L + 1;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 tw[L];
  const int y = blockIdx.x, chunk = blockIdx.y, Nx = g.Nx;
  const int Lr = RT ? 2 * Nx : L;
  load_tw<T2>(tw, twg, Lr);
  const T2* src = X + (((int64_t)chunk * g.Ny + y) * Lr) * NR;
  for (int i = threadIdx.x; i < NR * Lr; i += NT) {
    const int kx = i / NR, s = i - kx * NR;
    buf[s * LS + kx] = src[i];
  }
  __syncthreads();
  lds_fft_any<T2, L, NR, LS, NT, +1, RT>(buf, tw, g.px);
  for (int i = threadIdx.x; i < NR * Nx; i += NT) {
    const int s = i / Nx, x = i - s * Nx, cz = chunk * NR + s;
    if (cz < g.Nz3) {
      const int c = cz / g.Nz, z = cz - c * g.Nz;
      const int64_t ci = ((int64_t)z * g.Ny + y) * Nx + x;
      const int64_t vi = g.slot[c] * g.nc + ci;
      T2 d1 = vin[vi] * asiga[ci];
      T2 d2 = d1 - buf[s * LS + x] * dzn[z];
      vout[vi] = d2 * sq[z];
    }
  }
}

// ------------------------------------------ half-length FFT passes ----
// The x and y transforms are zero-padded: 2N points of which only N are
// non-zero going forward (x < Nx, y < Ny), and only N outputs are kept going
// back.  So each 2N-point transform is two N-point transforms:
//   forward  X[2k] = F_N(x)[k],  X[2k+1] = F_N(x w)[k],  w[n] = exp(-2 pi i n/2N)
//   inverse  x[n] = F_N^-1(X_even)[n] + conj(w[n]) F_N^-1(X_odd)[n],  n < N
// An LDS row then holds N points, so a block keeps twice the rows per LDS
// budget.  The row data of the second transform stay in registers.
template <typename R> constexpr int nr_half(int L) { return nr_for<R>(L / 2); }
// LDS row stride of the half-length passes.  Their row-transposed accesses
// (lane = row + NR * point, NR = 8 rows) reach rows s at stride LS: with
// 8-byte complex64 elements LS = N + 2 puts the sixteen lanes of one LDS
// cycle on sixteen distinct bank pairs (N + 1 paired rows s and points p
// with equal s + p; the complex64 passes measured 3-6% faster), with
// 16-byte FP64 elements an odd LS already spreads each cycle's eight lanes
template <typename R> constexpr int ls_half(int N) { return sizeof(R) == 4 ? N + 2 : N + 1; }

// The weights w[n] = exp(-2 pi i n/2N) are read from the global twiddle
// table in FP64 (37 KB of LDS per block instead of 41 KB: four blocks per CU
// instead of three; the four 512-point FP64 passes together 6.5% faster) and
// copied to LDS in complex64 (there the y pass measured 8% slower reading
// them from the global table).
template <typename R> constexpr bool wh_in_lds() { return sizeof(R) == 4; }

template <typename T2, int N, bool WH>
__device__ inline void load_tw_half(T2* twN, T2* wh, const T2* g) {
  for (int i = threadIdx.x; i < N; i += NT) {
    twN[i] = g[2 * i];  // exp(-2 pi i t/N)
    if constexpr (WH) wh[i] = g[i];
  }
}

template <typename R, int L>
__global__ __launch_bounds__(NT) void k_fwd_x_h(const cx_t<R>* __restrict__ v,
                                                const cx_t<R>* __restrict__ gsig,
                                                cx_t<R>* __restrict__ X,
                                                const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int N = L / 2, NR = nr_half<R>(L), LS = ls_half<R>(N), PER = (NR * N + NT - 1) / NT;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 twN[N], whs[wh_in_lds<R>() ? N : 1];
  const T2* __restrict__ wh = wh_in_lds<R>() ? whs : twg;
  const int y = blockIdx.x, chunk = x_chunk(g, blockIdx.y);  // Nx = N
  load_tw_half<T2, N, wh_in_lds<R>()>(twN, whs, twg);
  T2 xin[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    xin[q] = mk<T2>(0, 0);
    if (i < NR * N) {
      const int s = i / N, x = i - s * N, cz = chunk * NR + s;
      if (cz < g.Nz3) {
        const int c = cz / g.Nz, z = cz - c * g.Nz;
        const int64_t ci = ((int64_t)z * g.Ny + y) * N + x;
        xin[q] = v[g.slot[c] * g.nc + ci] * gsig[ci];
      }
      buf[s * LS + x] = xin[q];
    }
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, -1>(buf, twN);
  // X tile for this chunk: [y][kx][s].  The even kx stay in registers until
  // the odd ones are done, so the two halves of each stored pair (64 B each
  // for eight complex64 rows) reach the cache together and no line is
  // written twice
  T2* dst = X + (((int64_t)chunk * g.Ny + y) * L) * NR;
  T2 ev[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    ev[q] = mk<T2>(0, 0);
    if (i < NR * N) {
      const int k = i / NR, s = i - k * NR;
      ev[q] = buf[s * LS + k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    if (i < NR * N) {
      const int s = i / N, x = i - s * N;
      buf[s * LS + x] = xin[q] * wh[x];
    }
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, -1>(buf, twN);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    if (i < NR * N) {
      const int k = i / NR, s = i - k * NR;
      dst[(2 * k) * NR + s] = ev[q];
      dst[(2 * k + 1) * NR + s] = buf[s * LS + k];
    }
  }
}

template <typename R, int L>
__global__ __launch_bounds__(NT) void k_fwd_y_h(const cx_t<R>* __restrict__ X,
                                                cx_t<R>* __restrict__ S,
                                                const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int N = L / 2, NR = nr_half<R>(L), LS = ls_half<R>(N), PER = (NR * N + NT - 1) / NT;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 twN[N], whs[wh_in_lds<R>() ? N : 1];
  const T2* __restrict__ wh = wh_in_lds<R>() ? whs : twg;
  int kx, chunk;  // Ny = N
  y_tile<NR * sizeof(T2) < 128>(kx, chunk);
  if (kx == g.Nx) return;  // x-Nyquist column is never used
  load_tw_half<T2, N, wh_in_lds<R>()>(twN, whs, twg);
  T2 xin[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    xin[q] = mk<T2>(0, 0);
    if (i < NR * N) {
      const int yy = i / NR, s = i - yy * NR, cz = chunk * NR + s;
      if (cz < g.Nz3) xin[q] = X[xidx(g, cz, yy, kx)];
      buf[s * LS + yy] = xin[q];
    }
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, -1>(buf, twN);
  for (int i = threadIdx.x; i < NR * N; i += NT) {
    const int k = i / NR, s = i - k * NR, cz = chunk * NR + s;
    if (2 * k != N && cz < g.Nz3) S[((int64_t)kx * L + 2 * k) * g.Nz3 + cz] = buf[s * LS + k];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    if (i < NR * N) {
      const int yy = i / NR, s = i - yy * NR;
      buf[s * LS + yy] = xin[q] * wh[yy];
    }
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, -1>(buf, twN);
  for (int i = threadIdx.x; i < NR * N; i += NT) {
    const int k = i / NR, s = i - k * NR, cz = chunk * NR + s;
    if (cz < g.Nz3) S[((int64_t)kx * L + 2 * k + 1) * g.Nz3 + cz] = buf[s * LS + k];
  }
}

template <typename R, int L>
__global__ __launch_bounds__(NT) void k_inv_y_h(const cx_t<R>* __restrict__ S,
                                                cx_t<R>* __restrict__ X,
                                                const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int N = L / 2, NR = nr_half<R>(L), LS = ls_half<R>(N), PER = (NR * N + NT - 1) / NT;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 twN[N], whs[wh_in_lds<R>() ? N : 1];
  const T2* __restrict__ wh = wh_in_lds<R>() ? whs : twg;
  int kx, chunk;  // Ny = N
  y_tile<NR * sizeof(T2) < 128>(kx, chunk);
  if (kx == g.Nx) {  // x-Nyquist: zero (InitialTransposeRepack, Distributed_FFT.F90:723-724)
    for (int i = threadIdx.x; i < NR * N; i += NT) {
      const int yy = i / NR, s = i - yy * NR, cz = chunk * NR + s;
      if (cz < g.Nz3) X[xidx(g, cz, yy, kx)] = mk<T2>(0, 0);
    }
    return;
  }
  load_tw_half<T2, N, wh_in_lds<R>()>(twN, whs, twg);
  for (int i = threadIdx.x; i < NR * N; i += NT) {
    const int m = i / NR, s = i - m * NR, cz = chunk * NR + s;
    buf[s * LS + m] = (cz < g.Nz3) ? S[((int64_t)kx * L + 2 * m) * g.Nz3 + cz] : mk<T2>(0, 0);
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, +1>(buf, twN);
  T2 acc[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    acc[q] = mk<T2>(0, 0);
    if (i < NR * N) {
      const int yy = i / NR, s = i - yy * NR;
      acc[q] = buf[s * LS + yy];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NR * N; i += NT) {
    const int m = i / NR, s = i - m * NR, cz = chunk * NR + s;
    buf[s * LS + m] = (cz < g.Nz3) ? S[((int64_t)kx * L + 2 * m + 1) * g.Nz3 + cz] : mk<T2>(0, 0);
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, +1>(buf, twN);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    if (i < NR * N) {
      const int yy = i / NR, s = i - yy * NR, cz = chunk * NR + s;
      T2 w = wh[yy];
      w.y = -w.y;
      if (cz < g.Nz3) X[xidx(g, cz, yy, kx)] = acc[q] + buf[s * LS + yy] * w;
    }
  }
}

template <typename R, int L>
__global__ __launch_bounds__(NT) void k_inv_x_h(const cx_t<R>* __restrict__ X,
                                                const cx_t<R>* __restrict__ vin,
                                                const cx_t<R>* __restrict__ asiga,
                                                const R* __restrict__ sq,
                                                const R* __restrict__ dzn,
                                                cx_t<R>* __restrict__ vout,
                                                const cx_t<R>* __restrict__ twg, Geo g) {
  using T2 = cx_t<R>;
  constexpr int N = L / 2, NR = nr_half<R>(L), LS = ls_half<R>(N), PER = (NR * N + NT - 1) / NT;
  __shared__ T2 buf[NR * LS];
  __shared__ T2 twN[N], whs[wh_in_lds<R>() ? N : 1];
  const T2* __restrict__ wh = wh_in_lds<R>() ? whs : twg;
  const int y = blockIdx.x, chunk = x_chunk(g, blockIdx.y);  // Nx = N
  load_tw_half<T2, N, wh_in_lds<R>()>(twN, whs, twg);
  const T2* src = X + (((int64_t)chunk * g.Ny + y) * L) * NR;
  // even and odd kx of each stored pair are read together (the odd ones
  // wait in registers), so each line of X is fetched once
  T2 od[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    od[q] = mk<T2>(0, 0);
    if (i < NR * N) {
      const int m = i / NR, s = i - m * NR;
      buf[s * LS + m] = src[(2 * m) * NR + s];
      od[q] = src[(2 * m + 1) * NR + s];
    }
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, +1>(buf, twN);
  T2 acc[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    acc[q] = mk<T2>(0, 0);
    if (i < NR * N) {
      const int s = i / N, x = i - s * N;
      acc[q] = buf[s * LS + x];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    if (i < NR * N) {
      const int m = i / NR, s = i - m * NR;
      buf[s * LS + m] = od[q];
    }
  }
  __syncthreads();
  lds_fft<T2, N, NR, LS, NT, +1>(buf, twN);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int i = threadIdx.x + q * NT;
    if (i < NR * N) {
      const int s = i / N, x = i - s * N, cz = chunk * NR + s;
      if (cz < g.Nz3) {
        T2 w = wh[x];
        w.y = -w.y;
        const T2 f = acc[q] + buf[s * LS + x] * w;
        const int c = cz / g.Nz, z = cz - c * g.Nz;
        const int64_t ci = ((int64_t)z * g.Ny + y) * N + x;
        const int64_t vi = g.slot[c] * g.nc + ci;
        const T2 d1 = vin[vi] * asiga[ci];
        const T2 d2 = d1 - f * dzn[z];
        vout[vi] = d2 * sq[z];
      }
    }
  }
}

// ---------------------------------------------------------------- K3 ----
// One block per stored (kx,ky) quadrant block, serving the four spectrum
// points p = 0:(kx,ky) 1:(-kx,ky) 2:(kx,-ky) 3:(-kx,-ky) with
//   out_x = Sxx x + s m Sxy y + s Axz z
//   out_y = s m Sxy x + Syy y + m Ayz z
//   out_z = -s Axz^T x - m Ayz^T y + Szz z       (s = +-1 for +-kx, m for +-ky)
// (VERTICAL_MULT_GENERAL, apply_ie_operator_module.F90:276-292).  The six
// components (xx,xy,yy,zz packed symmetric; xz,yz full) stream through two
// LDS buffers: while the block works on component k, the coalesced global
// loads of component k+1 are in flight in registers.  Thread (row i, part)
// owns receiver layer i for all four points and sums source layers
// j = part, part+split, ...; partial sums are combined with lane shuffles.
// Register arrays are spelled out as named variables (no private arrays, so
// nothing is placed in scratch memory).
#define G2G_FOR4(X) X(0) X(1) X(2) X(3)
#define G2G_FOR16(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)

template <typename R, int PF>  // PF prefetch registers per thread: Nz*Nz <= PF*NT
__global__ __launch_bounds__(NT) void k_contract(cx_t<R>* __restrict__ S,
                                                 const cx_t<R>* __restrict__ G, Geo g,
                                                 int64_t BS, int split_log) {
  using T2 = cx_t<R>;
  static_assert(PF == 4 || PF == 16, "PF must be 4 or 16");
  extern __shared__ unsigned char smem_raw[];
  const int Nz = g.Nz, Nz3 = g.Nz3, NZP = Nz + 1;
  const int P = Nz * (Nz + 1) / 2;
  const int GB = Nz * NZP;                    // one LDS component buffer
  T2* vin = reinterpret_cast<T2*>(smem_raw);  // [4][3Nz]
  T2* gbuf = vin + 4 * Nz3;                   // [2][GB]
  const int b = blockIdx.x, kx = b % g.Nx, ky = b / g.Nx;
  const int Nx2 = 2 * g.Nx, Ny2 = 2 * g.Ny;
  // point p: bit 0 = -kx (x-Nyquist column when kx == 0), bit 1 = -ky
  const int gNx = g.Nx, gNy = g.Ny;
#define PX(p) (((p) & 1) ? (kx == 0 ? gNx : Nx2 - kx) : kx)
#define PY(p) (((p) & 2) ? (ky == 0 ? gNy : Ny2 - ky) : ky)
#define VAL(p) ((!((p) & 1) || kx > 0) && (!((p) & 2) || ky > 0))
  const T2* blk = G + (int64_t)b * BS;
  const T2 zero = mk<T2>(0, 0);
#define G2G_PFDECL(q) T2 pf##q = zero;
  G2G_FOR16(G2G_PFDECL)
#undef G2G_PFDECL
#define G2G_CLEN(c) ((c) < 4 ? P : Nz * Nz)
#define G2G_CSRC(c) ((c) < 4 ? blk + (c) * P : blk + 4 * P + ((c) - 4) * Nz * Nz)
#define G2G_LD(q) if ((q) < PF && threadIdx.x + (q) * NT < n_) pf##q = src_[threadIdx.x + (q) * NT];
#define G2G_PREFETCH(c)                  \
  {                                      \
    const T2* src_ = G2G_CSRC(c);        \
    const int n_ = G2G_CLEN(c);          \
    G2G_FOR16(G2G_LD)                    \
  }
#define G2G_ST(q)                                          \
  if ((q) < PF && threadIdx.x + (q) * NT < n_) {           \
    const int e = threadIdx.x + (q) * NT;                  \
    if (c_ < 4) dst_[e] = pf##q;                           \
    else dst_[(e / Nz) * NZP + e % Nz] = pf##q;            \
  }
#define G2G_STASH(c, dst)            \
  {                                  \
    const int c_ = (c);              \
    T2* dst_ = (dst);                \
    const int n_ = G2G_CLEN(c_);     \
    G2G_FOR16(G2G_ST)                \
  }
  G2G_PREFETCH(0)
  for (int i = threadIdx.x; i < 4 * Nz3; i += NT) {
    const int p = i / Nz3, e = i - p * Nz3;
    T2 vv = zero;
    if (VAL(p)) vv = S[((int64_t)PX(p) * Ny2 + PY(p)) * Nz3 + e];
    vin[i] = vv;
  }
  G2G_STASH(0, gbuf)
  __syncthreads();
  const int split = 1 << split_log;
  const int part = threadIdx.x & (split - 1);
  const int i = threadIdx.x >> split_log;
  const bool active = i < Nz;
  // sign of point p: s = -1 for -kx (bit 0), m = -1 for -ky (bit 1)
#define G2G_SG(p) (((p) & 1) ? R(-1) : R(1))
#define G2G_MR(p) (((p) & 2) ? R(-1) : R(1))
#define G2G_ACC(p) T2 ax##p = zero, ay##p = zero, az##p = zero;
  G2G_FOR4(G2G_ACC)
#undef G2G_ACC
  for (int comp = 0; comp < 6; comp++) {
    const T2* cur = gbuf + (comp & 1) * GB;
    if (comp < 5) G2G_PREFETCH(comp + 1)
    if (active) {
#define G2G_T(p) T2 t1##p = zero, t2##p = zero;
      G2G_FOR4(G2G_T)
#undef G2G_T
      if (comp < 4) {
        const int ii = i * (i + 1) / 2;
        const int off1 = comp == 0 ? 0 : (comp == 3 ? 2 * Nz : Nz);  // x for xx, z for zz, y otherwise
        for (int j = part; j < Nz; j += split) {
          const T2 sv = cur[j <= i ? j + ii : i + j * (j + 1) / 2];
#define G2G_SYM(p)                                      \
  {                                                     \
    const T2* v = vin + (p) * Nz3;                      \
    cmac(t1##p, sv, v[off1 + j]);                       \
    if (comp == 1) cmac(t2##p, sv, v[j]);               \
  }
          G2G_FOR4(G2G_SYM)
#undef G2G_SYM
        }
#define G2G_FOLD(p)                                                  \
  if (comp == 0) ax##p = ax##p + t1##p;                              \
  else if (comp == 1) {                                              \
    ax##p = ax##p + t1##p * (G2G_SG(p) * G2G_MR(p));                 \
    ay##p = ay##p + t2##p * (G2G_SG(p) * G2G_MR(p));                 \
  } else if (comp == 2) ay##p = ay##p + t1##p;                       \
  else az##p = az##p + t1##p;
        G2G_FOR4(G2G_FOLD)
#undef G2G_FOLD
      } else {
        const int offT = comp == 4 ? 0 : Nz;  // A^T acts on x (xz) or y (yz)
        for (int j = part; j < Nz; j += split) {
          const T2 a = cur[j * NZP + i];   // A(i,j)
          const T2 at = cur[i * NZP + j];  // A(j,i)
#define G2G_ASY(p)                          \
  {                                         \
    const T2* v = vin + (p) * Nz3;          \
    cmac(t1##p, a, v[2 * Nz + j]);          \
    cmac(t2##p, at, v[offT + j]);           \
  }
          G2G_FOR4(G2G_ASY)
#undef G2G_ASY
        }
#define G2G_FOLD(p)                                   \
  if (comp == 4) {                                    \
    ax##p = ax##p + t1##p * G2G_SG(p);                \
    az##p = az##p - t2##p * G2G_SG(p);                \
  } else {                                            \
    ay##p = ay##p + t1##p * G2G_MR(p);                \
    az##p = az##p - t2##p * G2G_MR(p);                \
  }
        G2G_FOR4(G2G_FOLD)
#undef G2G_FOLD
      }
    }
    if (comp < 5) G2G_STASH(comp + 1, gbuf + ((comp + 1) & 1) * GB)
    __syncthreads();
  }
  for (int o = 1; o < split; o <<= 1) {
#define G2G_RED(p)                                                       \
  ax##p.x += __shfl_xor(ax##p.x, o); ax##p.y += __shfl_xor(ax##p.y, o);  \
  ay##p.x += __shfl_xor(ay##p.x, o); ay##p.y += __shfl_xor(ay##p.y, o);  \
  az##p.x += __shfl_xor(az##p.x, o); az##p.y += __shfl_xor(az##p.y, o);
    G2G_FOR4(G2G_RED)
#undef G2G_RED
  }
  if (active && part == 0) {
#define G2G_OUT(p)                                                  \
  {                                                                 \
    T2* dst = S + ((int64_t)PX(p) * Ny2 + PY(p)) * Nz3;             \
    if (!VAL(p)) ax##p = ay##p = az##p = zero;                      \
    dst[i] = ax##p;                                                 \
    dst[Nz + i] = ay##p;                                            \
    dst[2 * Nz + i] = az##p;                                        \
  }
    G2G_FOR4(G2G_OUT)
#undef G2G_OUT
  }
#undef G2G_SG
#undef G2G_MR
#undef PX
#undef PY
#undef VAL
#undef G2G_CLEN
#undef G2G_CSRC
#undef G2G_LD
#undef G2G_PREFETCH
#undef G2G_ST
#undef G2G_STASH
}

// Layer counts above 64 (up to MAX_NZ): the products of k_contract with
// every tensor component staged through LDS in tiles of TJ source columns,
// which are contiguous in the packed-upper and in the column-major storage
// alike.  Thread (row i, part) owns receiver layer i of all four points.
// For a symmetric component S a tile of columns [j0, j1) gives row i the
// upper terms S(i,j) v(j), j >= i, from column j, and, when j0 <= i < j1,
// the lower terms S(k,i) v(k), k < i, from column i; for a full component A
// (xz, yz) row i takes A(i,j) v_z(j) of the tile's columns and, when i is in
// the tile, the transposed terms A(k,i) v(k) over all k from column i.  So
// every stored entry enters each product it belongs to exactly once.
template <typename R>
__global__ __launch_bounds__(NT) void k_contract_tiled(cx_t<R>* __restrict__ S,
                                                       const cx_t<R>* __restrict__ G, Geo g,
                                                       int64_t BS, int split_log, int TJ) {
  using T2 = cx_t<R>;
  extern __shared__ unsigned char smem_raw[];
  const int Nz = g.Nz, Nz3 = g.Nz3;
  const int64_t P = (int64_t)Nz * (Nz + 1) / 2;
  T2* vin = reinterpret_cast<T2*>(smem_raw);  // [4][3Nz]
  T2* tile = vin + 4 * Nz3;                   // at most TJ*Nz entries
  const int b = blockIdx.x, kx = b % g.Nx, ky = b / g.Nx;
  const int Nx2 = 2 * g.Nx, Ny2 = 2 * g.Ny;
  const int gNx = g.Nx, gNy = g.Ny;
#define PX(p) (((p) & 1) ? (kx == 0 ? gNx : Nx2 - kx) : kx)
#define PY(p) (((p) & 2) ? (ky == 0 ? gNy : Ny2 - ky) : ky)
#define VAL(p) ((!((p) & 1) || kx > 0) && (!((p) & 2) || ky > 0))
  const T2* blk = G + (int64_t)b * BS;
  const T2 zero = mk<T2>(0, 0);
  for (int t = threadIdx.x; t < 4 * Nz3; t += NT) {
    const int p = t / Nz3, e = t - p * Nz3;
    T2 vv = zero;
    if (VAL(p)) vv = S[((int64_t)PX(p) * Ny2 + PY(p)) * Nz3 + e];
    vin[t] = vv;
  }
  const int split = 1 << split_log;
  const int part = threadIdx.x & (split - 1);
  const int i = threadIdx.x >> split_log;
  const bool active = i < Nz;
#define G2G_SG(p) (((p) & 1) ? R(-1) : R(1))
#define G2G_MR(p) (((p) & 2) ? R(-1) : R(1))
#define G2G_ACC(p) T2 ax##p = zero, ay##p = zero, az##p = zero;
  G2G_FOR4(G2G_ACC)
#undef G2G_ACC
  for (int comp = 0; comp < 6; comp++) {
    const bool sym = comp < 4;
    const T2* src = sym ? blk + comp * P : blk + 4 * P + (int64_t)(comp - 4) * Nz * Nz;
    // t1: the component acting on its source (x for xx, y for xy and yy, z
    // for zz, xz and yz); t2: xy on x, or the transposed xz / yz
    const int off1 = comp == 0 ? 0 : (comp == 3 || comp >= 4 ? 2 * Nz : Nz);
    const int off2 = comp == 5 ? Nz : 0;
#define G2G_T(p) T2 t1##p = zero, t2##p = zero;
    G2G_FOR4(G2G_T)
#undef G2G_T
    for (int j0 = 0; j0 < Nz; j0 += TJ) {
      const int j1 = min(j0 + TJ, Nz);
      const int64_t base = sym ? (int64_t)j0 * (j0 + 1) / 2 : (int64_t)j0 * Nz;
      const int64_t len = sym ? (int64_t)j1 * (j1 + 1) / 2 - base : (int64_t)(j1 - j0) * Nz;
      __syncthreads();  // the previous tile (and, first time, vin) is consumed / written
      for (int64_t t = threadIdx.x; t < len; t += NT) tile[t] = ld_stream(src + base + t);
      __syncthreads();
      if (active) {
        if (sym) {
          for (int j = max(j0, i) + part; j < j1; j += split) {
            const T2 sv = tile[i + (int64_t)j * (j + 1) / 2 - base];
#define G2G_UP(p)                                     \
  {                                                   \
    const T2* v = vin + (p) * Nz3;                    \
    cmac(t1##p, sv, v[off1 + j]);                     \
    if (comp == 1) cmac(t2##p, sv, v[j]);             \
  }
            G2G_FOR4(G2G_UP)
#undef G2G_UP
          }
          if (i >= j0 && i < j1) {
            const T2* col = tile + (int64_t)i * (i + 1) / 2 - base;
            for (int k = part; k < i; k += split) {
              const T2 sv = col[k];
#define G2G_LO(p)                                     \
  {                                                   \
    const T2* v = vin + (p) * Nz3;                    \
    cmac(t1##p, sv, v[off1 + k]);                     \
    if (comp == 1) cmac(t2##p, sv, v[k]);             \
  }
              G2G_FOR4(G2G_LO)
#undef G2G_LO
            }
          }
        } else {
          for (int j = j0 + part; j < j1; j += split) {
            const T2 a = tile[(int64_t)(j - j0) * Nz + i];  // A(i,j)
#define G2G_AF(p) cmac(t1##p, a, vin[(p) * Nz3 + off1 + j]);
            G2G_FOR4(G2G_AF)
#undef G2G_AF
          }
          if (i >= j0 && i < j1) {
            const T2* col = tile + (int64_t)(i - j0) * Nz;
            for (int k = part; k < Nz; k += split) {
              const T2 at = col[k];  // A(k,i)
#define G2G_AT(p) cmac(t2##p, at, vin[(p) * Nz3 + off2 + k]);
              G2G_FOR4(G2G_AT)
#undef G2G_AT
            }
          }
        }
      }
    }
#define G2G_FOLD(p)                                                  \
  if (comp == 0) ax##p = ax##p + t1##p;                              \
  else if (comp == 1) {                                              \
    ax##p = ax##p + t1##p * (G2G_SG(p) * G2G_MR(p));                 \
    ay##p = ay##p + t2##p * (G2G_SG(p) * G2G_MR(p));                 \
  } else if (comp == 2) ay##p = ay##p + t1##p;                       \
  else if (comp == 3) az##p = az##p + t1##p;                         \
  else if (comp == 4) {                                              \
    ax##p = ax##p + t1##p * G2G_SG(p);                               \
    az##p = az##p - t2##p * G2G_SG(p);                               \
  } else {                                                           \
    ay##p = ay##p + t1##p * G2G_MR(p);                               \
    az##p = az##p - t2##p * G2G_MR(p);                               \
  }
    G2G_FOR4(G2G_FOLD)
#undef G2G_FOLD
  }
  for (int o = 1; o < split; o <<= 1) {
#define G2G_RED(p)                                                       \
  ax##p.x += __shfl_xor(ax##p.x, o); ax##p.y += __shfl_xor(ax##p.y, o);  \
  ay##p.x += __shfl_xor(ay##p.x, o); ay##p.y += __shfl_xor(ay##p.y, o);  \
  az##p.x += __shfl_xor(az##p.x, o); az##p.y += __shfl_xor(az##p.y, o);
    G2G_FOR4(G2G_RED)
#undef G2G_RED
  }
  if (active && part == 0) {
#define G2G_OUT(p)                                                  \
  {                                                                 \
    T2* dst = S + ((int64_t)PX(p) * Ny2 + PY(p)) * Nz3;             \
    if (!VAL(p)) ax##p = ay##p = az##p = zero;                      \
    dst[i] = ax##p;                                                 \
    dst[Nz + i] = ay##p;                                            \
    dst[2 * Nz + i] = az##p;                                        \
  }
    G2G_FOR4(G2G_OUT)
#undef G2G_OUT
  }
#undef G2G_SG
#undef G2G_MR
#undef PX
#undef PY
#undef VAL
}

// Two-deep variant of k_contract: two tensor components in flight in
// registers while a third is computed from LDS.
// NZT > 0: Nz known at compile time (a power of two dividing NTC, with
// Nz >= NTC/Nz lanes per row), so the
// per-row loops over source layers unroll and their LDS reads issue
// together; NZT = 0: any Nz.  The inputs are stored pre-multiplied by the
// point signs D_p = diag(s, m, 1), so every product accumulates straight
// into the outputs (out_p = D_p G0 (D_p v_p), the sign pattern of
// VERTICAL_MULT_GENERAL) and the signs are applied once at the end.
template <typename R, int PF, int NTC, int NZT>  // NTC threads; Nz*Nz <= PF*NTC
__global__ __launch_bounds__(NTC) void k_contract2(cx_t<R>* __restrict__ S,
                                                 const cx_t<R>* __restrict__ G, Geo g,
                                                 int64_t BS, int split_log) {
  using T2 = cx_t<R>;
  static_assert(PF == 4 || PF == 8 || PF == 16, "PF must be 4, 8 or 16");
  // compile-time Nz: a power of two with at least one source layer per lane
  static_assert(NZT == 0 || (NTC % NZT == 0 && (NZT & (NZT - 1)) == 0 && NZT * NZT >= NTC), "NZT");
  extern __shared__ unsigned char smem_raw[];
  const int Nz = NZT > 0 ? NZT : g.Nz;
  const int Nz3 = 3 * Nz, NZP = Nz + 1;
  const int P = Nz * (Nz + 1) / 2;
  const int GB = Nz * NZP;                    // one LDS component buffer
  T2* vin = reinterpret_cast<T2*>(smem_raw);  // [4][3Nz]
  T2* gbuf = vin + 4 * Nz3;                   // [2][GB]
  const int b = blockIdx.x, kx = b % g.Nx, ky = b / g.Nx;
  const int Nx2 = 2 * g.Nx, Ny2 = 2 * g.Ny;
  // point p: bit 0 = -kx (x-Nyquist column when kx == 0), bit 1 = -ky
  const int gNx = g.Nx, gNy = g.Ny;
#define PX(p) (((p) & 1) ? (kx == 0 ? gNx : Nx2 - kx) : kx)
#define PY(p) (((p) & 2) ? (ky == 0 ? gNy : Ny2 - ky) : ky)
#define VAL(p) ((!((p) & 1) || kx > 0) && (!((p) & 2) || ky > 0))
  const T2* blk = G + (int64_t)b * BS;
  const T2 zero = mk<T2>(0, 0);
#define G2G_PFDECL(q) T2 pf##q = zero, pg##q = zero;
  G2G_FOR16(G2G_PFDECL)
#undef G2G_PFDECL
#define G2G_CLEN(c) ((c) < 4 ? P : Nz * Nz)
#define G2G_CSRC(c) ((c) < 4 ? blk + (c) * P : blk + 4 * P + ((c) - 4) * Nz * Nz)
  // loads are unconditional (index clamped into the component) so that the
  // compiler counts outstanding loads exactly and keeps the second set in flight
#define G2G_LDF(q) if ((q) < PF) pf##q = ld_stream(src_ + min((int)threadIdx.x + (q) * NTC, n_ - 1));
#define G2G_LDG(q) if ((q) < PF) pg##q = ld_stream(src_ + min((int)threadIdx.x + (q) * NTC, n_ - 1));
#define G2G_PREFETCH(c, SET)             \
  {                                      \
    const T2* src_ = G2G_CSRC(c);        \
    const int n_ = G2G_CLEN(c);          \
    G2G_FOR16(G2G_LD##SET)               \
  }
#define G2G_STX(q, r)                                      \
  if ((q) < PF && threadIdx.x + (q) * NTC < n_) {           \
    const int e = threadIdx.x + (q) * NTC;                  \
    if (c_ < 4) dst_[e] = r##q;                            \
    else dst_[(e / Nz) * NZP + e % Nz] = r##q;             \
  }
#define G2G_STF(q) G2G_STX(q, pf)
#define G2G_STG(q) G2G_STX(q, pg)
#define G2G_STASH(c, SET, dst)       \
  {                                  \
    const int c_ = (c);              \
    T2* dst_ = (dst);                \
    const int n_ = G2G_CLEN(c_);     \
    G2G_FOR16(G2G_ST##SET)           \
  }
  G2G_PREFETCH(0, F)
  G2G_PREFETCH(1, G)
  for (int i = threadIdx.x; i < 4 * Nz3; i += NTC) {
    const int p = i / Nz3, e = i - p * Nz3;
    T2 vv = zero;
    if (VAL(p)) vv = S[((int64_t)PX(p) * Ny2 + PY(p)) * Nz3 + e];
    // D_p: x times s (-1 for -kx), y times m (-1 for -ky)
    if ((e < Nz && (p & 1)) || (e >= Nz && e < 2 * Nz && (p & 2))) vv = -vv;
    vin[i] = vv;
  }
  G2G_STASH(0, F, gbuf)
  __syncthreads();
  constexpr int SPLIT = NZT > 0 ? NTC / NZT : 1;
  const int split = NZT > 0 ? SPLIT : 1 << split_log;
  const int part = threadIdx.x & (split - 1);
  const int i = NZT > 0 ? (int)threadIdx.x / SPLIT : (int)threadIdx.x >> split_log;
  const bool active = i < Nz;
  const int nj = NZT > 0 ? NZT / SPLIT : (Nz - part + split - 1) / split;  // source layers per lane
#define G2G_ACC(p) T2 ax##p = zero, ay##p = zero, az##p = zero;
  G2G_FOR4(G2G_ACC)
#undef G2G_ACC
#pragma unroll
  for (int comp = 0; comp < 6; comp++) {
    const T2* cur = gbuf + (comp & 1) * GB;
    // registers: F holds components 0, 2, 4 and G 1, 3, 5; component
    // comp+2 is requested while comp is computed and comp+1 is in flight
    if (comp + 2 < 6) {
      if (comp & 1) G2G_PREFETCH(comp + 2, G)
      else G2G_PREFETCH(comp + 2, F)
    }
    if (active) {
      if (comp < 4) {
        const int ii = i * (i + 1) / 2;
        const int off1 = comp == 0 ? 0 : (comp == 3 ? 2 * Nz : Nz);  // x for xx, z for zz, y otherwise
#pragma unroll
        for (int jj = 0; jj < nj; jj++) {
          const int j = part + jj * split;
          const T2 sv = cur[j <= i ? j + ii : i + j * (j + 1) / 2];
#define G2G_SYM(p)                                      \
  {                                                     \
    const T2* v = vin + (p) * Nz3;                      \
    if (comp == 0) cmac(ax##p, sv, v[j]);               \
    else if (comp == 1) {                               \
      cmac(ax##p, sv, v[Nz + j]);                       \
      cmac(ay##p, sv, v[j]);                            \
    } else if (comp == 2) cmac(ay##p, sv, v[off1 + j]); \
    else cmac(az##p, sv, v[off1 + j]);                  \
  }
          G2G_FOR4(G2G_SYM)
#undef G2G_SYM
        }
      } else {
        const int offT = comp == 4 ? 0 : Nz;  // A^T acts on x (xz) or y (yz)
#pragma unroll
        for (int jj = 0; jj < nj; jj++) {
          const int j = part + jj * split;
          const T2 a = cur[j * NZP + i];     // A(i,j)
          const T2 atn = -cur[i * NZP + j];  // -A(j,i)
#define G2G_ASY(p)                                  \
  {                                                 \
    const T2* v = vin + (p) * Nz3;                  \
    if (comp == 4) cmac(ax##p, a, v[2 * Nz + j]);   \
    else cmac(ay##p, a, v[2 * Nz + j]);             \
    cmac(az##p, atn, v[offT + j]);                  \
  }
          G2G_FOR4(G2G_ASY)
#undef G2G_ASY
        }
      }
    }
    if (comp < 5) {
      if (comp & 1) G2G_STASH(comp + 1, F, gbuf + ((comp + 1) & 1) * GB)
      else G2G_STASH(comp + 1, G, gbuf + ((comp + 1) & 1) * GB)
    }
    __syncthreads();
  }
  for (int o = 1; o < split; o <<= 1) {
#define G2G_RED(p)                                                       \
  ax##p.x += __shfl_xor(ax##p.x, o); ax##p.y += __shfl_xor(ax##p.y, o);  \
  ay##p.x += __shfl_xor(ay##p.x, o); ay##p.y += __shfl_xor(ay##p.y, o);  \
  az##p.x += __shfl_xor(az##p.x, o); az##p.y += __shfl_xor(az##p.y, o);
    G2G_FOR4(G2G_RED)
#undef G2G_RED
  }
  if (active && part == 0) {
#define G2G_OUT(p)                                                  \
  {                                                                 \
    T2* dst = S + ((int64_t)PX(p) * Ny2 + PY(p)) * Nz3;             \
    if (!VAL(p)) ax##p = ay##p = az##p = zero;                      \
    dst[i] = ((p) & 1) ? -ax##p : ax##p;                            \
    dst[Nz + i] = ((p) & 2) ? -ay##p : ay##p;                       \
    dst[2 * Nz + i] = az##p;                                        \
  }
    G2G_FOR4(G2G_OUT)
#undef G2G_OUT
  }
#undef PX
#undef PY
#undef VAL
#undef G2G_CLEN
#undef G2G_CSRC
#undef G2G_LDF
#undef G2G_LDG
#undef G2G_PREFETCH
#undef G2G_STX
#undef G2G_STF
#undef G2G_STG
#undef G2G_STASH
}

// ------------------------------------------------------ K3 on MFMA ----
// The same product as k_contract2, as matrix-core tiles.  With the point
// signs s (for -kx) and m (for -ky), every point uses the same 3Nz x 3Nz
// block G0 = [[Sxx, Sxy, Axz], [Sxy, Syy, Ayz], [-Axz^T, -Ayz^T, Szz]]:
//   out_p = D_p G0 (D_p v_p),  D_p = diag(s, m, 1) per component
// (the sign pattern of VERTICAL_MULT_GENERAL, apply_ie_operator_module.F90:
// 276-292).  The four points form the columns of a real 16-column operand
// [Re v | Im v] (Gr pass) and [-Im v | Re v] (Gi pass) so that columns 0-3
// accumulate Re(G0 v_p) and columns 4-7 Im(G0 v_p) in one
// v_mfma_{f64,f32}_16x16x4 tile; columns 8-15 carry a second right-hand
// side (NRHS = 2: two vectors read the tensor once) or repeat the first.
// Lane l supplies A[row l%16][k l/16] and B[k l/16][col l%16] and receives
// four rows of column l%16 (mma_row).
template <typename R> struct v4t;
template <> struct v4t<double> { typedef double t __attribute__((ext_vector_type(4))); };
template <> struct v4t<float> { typedef float t __attribute__((ext_vector_type(4))); };
__device__ inline v4t<double>::t mma4(double a, double b, v4t<double>::t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4t<float>::t mma4(float a, float b, v4t<float>::t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// result row of accumulator element q in lane group kq = lane/16: the FP32
// tile holds rows 4*kq + q, the FP64 tile rows kq + 4*q (both measured)
template <typename R> __device__ inline int mma_row(int kq, int q);
template <> __device__ inline int mma_row<float>(int kq, int q) { return 4 * kq + q; }
template <> __device__ inline int mma_row<double>(int kq, int q) { return kq + 4 * q; }

// The work of every tensor component is spread over all NW waves: wave w
// takes the source layers [w Nz/NW, (w+1) Nz/NW) of every output tile, so
// each component keeps every wave's matrix cores busy.  The point operand is
// loaded once into registers (three input components x Nz/(4 NW) steps per lane);
// the tensor operand is read from LDS in its stored layout: packed symmetric
// components addressed directly (S(i,j) = S(j,i)), full components with column
// stride Nz+1.  Each wave accumulates all 3Nz/16 output tiles over its layers;
// the partial tiles are summed through LDS by a fixed halving tree (NW = 4:
// (w0 + w2) + (w1 + w3)), so results do not depend on scheduling.  DEEP = 2
// keeps two components in flight in registers, DEEP = 1 one (FP64: fewer
// registers, more waves per SIMD).
template <typename R, int NZ, int NRHS, int NW>
struct C3Cfg {
  static constexpr int NTH = 64 * NW, NZ3 = 3 * NZ, NT16 = NZ / 16, KW = NZ / NW, KS = KW / 4;
  static constexpr int NG = NRHS == 1 ? 1 : NRHS / 2;  // 16-column operand groups
  static constexpr int P = NZ * (NZ + 1) / 2, FULL = NZ * NZ, SF = NZ + 1, GBS = NZ * SF;
  static constexpr int PF = (FULL + NTH - 1) / NTH;
  static constexpr int SL = 3 * NT16 * 256;  // reals per partial-tile slot
  static constexpr size_t vin_bytes = sizeof(cx_t<R>) * NRHS * 4 * NZ3;
  static constexpr size_t stage_bytes = sizeof(cx_t<R>) * 2 * GBS;
  static constexpr size_t red_bytes = sizeof(R) * (NW / 2) * SL;
  static constexpr size_t smem = vin_bytes + (stage_bytes > red_bytes ? stage_bytes : red_bytes);
};

// MIR (NRHS = 2 only): the second right-hand side is the x-y mirror of the
// first, as in k_contract1: blocks cover kx <= ky, and "rhs 1" is the quad
// (ky, kx), read at the transposed coordinates (PY(p), PX(p)) with x <-> y
// exchanged and stored the same way, with the signs of point p; a diagonal
// block (kx = ky) leaves it out.  One-vector products then fill the tile
// columns that would otherwise repeat, from one tensor read per pair.
template <typename R, int NZ, int NRHS, int DEEP, int NW, bool PAD, bool MIR = false>
__device__ __forceinline__ void contract3_body(cx_t<R>* __restrict__ S, int64_t sst,
                                               const cx_t<R>* __restrict__ G, const Geo& g, int64_t BS) {
  using T2 = cx_t<R>;
  using V4 = typename v4t<R>::t;
  using C = C3Cfg<R, NZ, NRHS, NW>;
  constexpr int NTH = C::NTH, NZ3 = C::NZ3, NT16 = C::NT16, KW = C::KW, KS = C::KS;
  constexpr int P = C::P, FULL = C::FULL, SF = C::SF, GBS = C::GBS, PF = C::PF, SL = C::SL, NG = C::NG;
  static_assert(NZ % 16 == 0 && PF <= 16, "Nz must be a multiple of 16, at most 64");
  static_assert((NW == 4 || NW == 8) && KW % 4 == 0, "four or eight waves, whole MFMA steps each");
  static_assert(NRHS == 1 || NRHS == 2 || NRHS == 4 || NRHS == 8, "1, 2, 4 or 8 right-hand sides");
  static_assert(DEEP >= 1 && DEEP <= 6, "one to six components in flight");
  extern __shared__ unsigned char smem_raw[];
  T2* vin = reinterpret_cast<T2*>(smem_raw);  // [NRHS][4][3NZ], point-signed inputs; rhs r at S + r*sst
  T2* gbuf = vin + NRHS * 4 * NZ3;            // [2][GBS]
  // PAD: the operator has nz < NZ layers; rows and columns nz..NZ-1 of every
  // tile are zero (zero point operand, zeroed or masked tensor entries, rows
  // not stored)
  const int nz = PAD ? g.Nz : NZ, nz3 = 3 * nz;
  const int Pn = PAD ? nz * (nz + 1) / 2 : P, Fn = PAD ? nz * nz : FULL;
  // x / nz and x / 3nz as a multiply and a shift: exact for x * divisor < 2^20
  // (here x < 3072 and divisor <= 192)
  const uint32_t mnz = PAD ? ((1u << 20) + nz - 1) / nz : 0, mnz3 = PAD ? ((1u << 20) + nz3 - 1) / nz3 : 0;
#define C3_DIVNZ(x) (PAD ? (int)(((uint32_t)(x) * mnz) >> 20) : (x) / NZ)
#define C3_DIVNZ3(x) (PAD ? (int)(((uint32_t)(x) * mnz3) >> 20) : (x) / NZ3)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  static_assert(!MIR || NRHS == 2, "the mirror takes the second operand group");
  // blocks walk ky fastest: concurrent blocks touch neighbouring spectrum
  // rows (measured 2-3% faster at 256x256 than kx fastest); with MIR block
  // t = ky (ky + 1) / 2 + kx, kx <= ky
  int kx, ky;
  if constexpr (MIR) {
    const int t = blockIdx.x;
    int y = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((y + 1) * (y + 2) / 2 <= t) y++;
    while (y * (y + 1) / 2 > t) y--;
    ky = y;
    kx = t - y * (y + 1) / 2;
  } else {
    kx = blockIdx.x / g.Ny;
    ky = blockIdx.x % g.Ny;
  }
  const int b = kx + g.Nx * ky;
  const bool two = !MIR || kx != ky;  // the mirrored set exists off the diagonal
  const int Nx2 = 2 * g.Nx, Ny2 = 2 * g.Ny, gNx = g.Nx, gNy = g.Ny;
#define PX(p) (((p) & 1) ? (kx == 0 ? gNx : Nx2 - kx) : kx)
#define PY(p) (((p) & 2) ? (ky == 0 ? gNy : Ny2 - ky) : ky)
#define VAL(p) ((!((p) & 1) || kx > 0) && (!((p) & 2) || ky > 0))
  // rhs r of point p: its spectrum row and the component stored as ci
#define C3_MSET(r) (MIR && (r) == 1)
#define C3_ROWP(r, p)                                                              \
  (C3_MSET(r) ? S + ((int64_t)PY(p) * Ny2 + PX(p)) * nz3                            \
              : S + (MIR ? 0 : (r) * sst) + ((int64_t)PX(p) * Ny2 + PY(p)) * nz3)
#define C3_CS(r, ci) ((C3_MSET(r) && (ci) < 2) ? 1 - (ci) : (ci))
#define C3_LIVE(r, p) (VAL(p) && (!C3_MSET(r) || two))
  const T2* blk = G + (int64_t)b * BS;
  const T2 zero = mk<T2>(0, 0);
  // DEEP register sets F, G, H, I, J, K; component c travels in set c % DEEP
  // (named registers: arrays of FP64 pairs were placed in scratch memory by
  // the compiler)
#define C3_DECL(q) T2 rf##q = zero, rg##q = zero, rh##q = zero, ri##q = zero, rj##q = zero, rk##q = zero;
  G2G_FOR16(C3_DECL)
#undef C3_DECL
  // loads are unconditional (index clamped into the component) so that the
  // compiler counts outstanding loads exactly and keeps later sets in flight
#define C3_LDF(q) if ((q) < PF) rf##q = ld_stream(src_ + min(tid + (q) * NTH, n_ - 1));
#define C3_LDG(q) if ((q) < PF) rg##q = ld_stream(src_ + min(tid + (q) * NTH, n_ - 1));
#define C3_LDH(q) if ((q) < PF) rh##q = ld_stream(src_ + min(tid + (q) * NTH, n_ - 1));
#define C3_LDI(q) if ((q) < PF) ri##q = ld_stream(src_ + min(tid + (q) * NTH, n_ - 1));
#define C3_LDJ(q) if ((q) < PF) rj##q = ld_stream(src_ + min(tid + (q) * NTH, n_ - 1));
#define C3_LDK(q) if ((q) < PF) rk##q = ld_stream(src_ + min(tid + (q) * NTH, n_ - 1));
#define C3_PREFETCH(c, SET)                                                           \
  {                                                                                   \
    const T2* src_ = (c) < 4 ? blk + (c) * Pn : blk + 4 * Pn + ((c) - 4) * Fn;        \
    const int n_ = (c) < 4 ? Pn : Fn;                                                 \
    G2G_FOR16(C3_LD##SET)                                                             \
  }
  // packed symmetric components as stored, full ones with column stride SF
#define C3_STX(q, r)                                                 \
  if ((q) < PF && tid + (q) * NTH < n_) {                            \
    const int e = tid + (q) * NTH, col = C3_DIVNZ(e);                \
    dst_[c_ < 4 ? e : col * SF + e - col * nz] = r##q;               \
  }
#define C3_STF(q) C3_STX(q, rf)
#define C3_STG(q) C3_STX(q, rg)
#define C3_STH(q) C3_STX(q, rh)
#define C3_STI(q) C3_STX(q, ri)
#define C3_STJ(q) C3_STX(q, rj)
#define C3_STK(q) C3_STX(q, rk)
#define C3_STASH(c, SET, dst)          \
  {                                    \
    const int c_ = (c);                \
    T2* dst_ = (dst);                  \
    const int n_ = c_ < 4 ? Pn : Fn;   \
    G2G_FOR16(C3_ST##SET)              \
  }
  // the same with the set given by its compile-time number
#define C3_SETS(OP, c, k, ...)                          \
  {                                                     \
    if constexpr ((k) == 0) OP(c, F, ##__VA_ARGS__)     \
    else if constexpr ((k) == 1) OP(c, G, ##__VA_ARGS__) \
    else if constexpr ((k) == 2) OP(c, H, ##__VA_ARGS__) \
    else if constexpr ((k) == 3) OP(c, I, ##__VA_ARGS__) \
    else if constexpr ((k) == 4) OP(c, J, ##__VA_ARGS__) \
    else OP(c, K, ##__VA_ARGS__)                        \
  }
  // the point rows: complex64 issues them first as unrolled clamped loads
  // (paired Nz = 64: 6% faster); FP64 keeps the plain loop (the early
  // loads were 20% slower there)
  constexpr int NV = NRHS * 4 * NZ3, NVQ = (NV + NTH - 1) / NTH;
  constexpr bool EARLY = sizeof(R) == 4;
  static_assert(!EARLY || NVQ <= 16, "point rows: at most sixteen registers per lane");
  // named registers (a private array was placed in scratch memory for some sizes)
#define C3_VQD(k) T2 vq##k = zero;
  G2G_FOR16(C3_VQD)
#undef C3_VQD
  // padded slot i of a point row: component e / NZ, layer e % NZ
#define C3_VQL(k)                                                                          \
  if ((k) < NVQ) {                                                                         \
    const int i = min(tid + (k) * NTH, NV - 1);                                            \
    const int rp = i / NZ3, e = i - rp * NZ3, p = rp & 3, ci = e / NZ, z = min(e % NZ, nz - 1); \
    vq##k = C3_ROWP(rp >> 2, p)[C3_CS(rp >> 2, ci) * nz + z];                               \
  }
#define C3_VQP(k)                                                                  \
  if ((k) < NVQ && tid + (k) * NTH < NV) {                                       \
    const int i = tid + (k) * NTH;                                               \
    const int rp = i / NZ3, e = i - rp * NZ3, p = rp & 3;                        \
    T2 vv = (C3_LIVE(rp >> 2, p) && e % NZ < nz) ? vq##k : zero;                 \
    /* D_p: x times s (-1 for -kx), y times m (-1 for -ky) */                    \
    if ((e < NZ && (p & 1)) || (e >= NZ && e < 2 * NZ && (p & 2))) vv = -vv;     \
    vin[i] = vv;                                                                 \
  }
  if constexpr (EARLY) {
    G2G_FOR16(C3_VQL)
  }
  C3_PREFETCH(0, F)
  if constexpr (DEEP > 1) C3_PREFETCH(1, G)
  if constexpr (DEEP > 2) C3_PREFETCH(2, H)
  if constexpr (DEEP > 3) C3_PREFETCH(3, I)
  if constexpr (DEEP > 4) C3_PREFETCH(4, J)
  if constexpr (DEEP > 5) C3_PREFETCH(5, K)
  if constexpr (EARLY) {
    G2G_FOR16(C3_VQP)
  } else {
    for (int i = tid; i < NV; i += NTH) {
      const int rp = i / NZ3, e = i - rp * NZ3, p = rp & 3, ci = e / NZ, z = e % NZ;
      T2 vv = zero;
      if (C3_LIVE(rp >> 2, p) && z < nz) vv = C3_ROWP(rp >> 2, p)[C3_CS(rp >> 2, ci) * nz + z];
      if ((e < NZ && (p & 1)) || (e >= NZ && e < 2 * NZ && (p & 2))) vv = -vv;
      vin[i] = vv;
    }
  }
  if constexpr (PAD) {
    // packed entries Pn..P-1 of both buffers (rows or columns past nz of the
    // symmetric components) stay zero: their stashes write 0..Pn-1 only
    for (int i = Pn + tid; i < P; i += NTH) gbuf[i] = gbuf[GBS + i] = zero;
  }
  C3_STASH(0, F, gbuf)
  __syncthreads();
  // operand roles: lane l supplies A[row l%16][k l/16] and B[k l/16][col l%16];
  // column = 4*part + point, part = 2*(rhs within the group) + (0 re | 1 im);
  // operand group q serves right-hand sides 2q and 2q+1
  const int rr = lane & 15, kq = lane >> 4;
  const int pnt = rr & 3, part = rr >> 2, rl = NRHS == 1 ? 0 : (part >> 1), reim = part & 1;
  const int jw = wave * KW;
  // the point operand of all three input components, loaded once
  R bR[NG][3][KS], bI[NG][3][KS];
#pragma unroll
  for (int q = 0; q < NG; q++)
#pragma unroll
    for (int ci = 0; ci < 3; ci++)
#pragma unroll
      for (int s = 0; s < KS; s++) {
        const T2 v = vin[((2 * q + rl) * 4 + pnt) * NZ3 + ci * NZ + jw + 4 * s + kq];
        bR[q][ci][s] = reim ? v.y : v.x;   // Gr pass: [Re v | Im v]
        bI[q][ci][s] = reim ? v.x : -v.y;  // Gi pass: [-Im v | Re v]
      }
  V4 acc[NG][3][NT16];
#pragma unroll
  for (int q = 0; q < NG; q++)
#pragma unroll
    for (int co = 0; co < 3; co++)
#pragma unroll
      for (int t = 0; t < NT16; t++) acc[q][co][t] = V4{0, 0, 0, 0};
  // one tensor component per step; the component is a compile-time value so
  // that accumulator and register-set indices stay static
  auto step = [&](auto CC) {
    constexpr int c = decltype(CC)::value;
    // the set of component c was stashed at the end of the previous step
    if constexpr (c + DEEP < 6) C3_SETS(C3_PREFETCH, c + DEEP, c % DEEP)
    const T2* cur = gbuf + (c & 1) * GBS;
    // input components of the step's first and second use
    constexpr int ciA = c == 0 ? 0 : (c == 3 || c >= 4 ? 2 : 1);
    constexpr int ciB = c == 1 ? 0 : (c >= 4 ? c - 4 : 0);
#pragma unroll
    for (int s = 0; s < KS; s++) {
      const int j = jw + 4 * s + kq;
#pragma unroll
      for (int t = 0; t < NT16; t++) {
        const int r = 16 * t + rr;
        if (c < 4) {
          // symmetric, packed by rows of the lower triangle (PAD: columns
          // j >= nz read the zeroed entries, rows r >= nz are not stored)
          const T2 a = cur[j <= r ? j + r * (r + 1) / 2 : r + j * (j + 1) / 2];
          // xx: x <- x; xy: x <- y (and y <- x below); yy: y <- y; zz: z <- z
          constexpr int co = c <= 1 ? 0 : (c == 2 ? 1 : 2);
#pragma unroll
          for (int q = 0; q < NG; q++) {
            acc[q][co][t] = mma4(a.x, bR[q][ciA][s], acc[q][co][t]);
            acc[q][co][t] = mma4(a.y, bI[q][ciA][s], acc[q][co][t]);
            if constexpr (c == 1) {  // S_xy also maps x to y
              acc[q][1][t] = mma4(a.x, bR[q][ciB][s], acc[q][1][t]);
              acc[q][1][t] = mma4(a.y, bI[q][ciB][s], acc[q][1][t]);
            }
          }
        } else {
          // A (xz or yz) at column j, row r; its use in z is -A^T
          constexpr int co = c - 4;
          T2 a = cur[j * SF + r];
          T2 at = cur[r * SF + j];
          if (PAD && j >= nz) a = at = zero;
#pragma unroll
          for (int q = 0; q < NG; q++) {
            acc[q][co][t] = mma4(a.x, bR[q][ciA][s], acc[q][co][t]);
            acc[q][co][t] = mma4(a.y, bI[q][ciA][s], acc[q][co][t]);
            acc[q][2][t] = mma4(-at.x, bR[q][ciB][s], acc[q][2][t]);
            acc[q][2][t] = mma4(-at.y, bI[q][ciB][s], acc[q][2][t]);
          }
        }
      }
    }
    if constexpr (c < 5) {
      C3_SETS(C3_STASH, c + 1, (c + 1) % DEEP, gbuf + ((c + 1) & 1) * GBS)
    }
    __syncthreads();
  };
  step(std::integral_constant<int, 0>{});
  step(std::integral_constant<int, 1>{});
  step(std::integral_constant<int, 2>{});
  step(std::integral_constant<int, 3>{});
  step(std::integral_constant<int, 4>{});
  step(std::integral_constant<int, 5>{});
  // partial tiles: slot [tile][q][lane]; halving tree, wave w adds wave w + h
  // one operand group at a time (the slots are reused; wave 0 finishes
  // reading a group before the next group is written)
  R* red = reinterpret_cast<R*>(gbuf);
  R* vo = reinterpret_cast<R*>(vin);  // results with the output signs, over vin
  const R sgx = (pnt & 1) ? R(-1) : R(1), sgy = (pnt & 2) ? R(-1) : R(1);
#pragma unroll
  for (int gq = 0; gq < NG; gq++) {
    auto put = [&](int slot) {
#pragma unroll
      for (int co = 0; co < 3; co++)
#pragma unroll
        for (int t = 0; t < NT16; t++)
#pragma unroll
          for (int q = 0; q < 4; q++) red[slot * SL + ((co * NT16 + t) * 4 + q) * 64 + lane] = acc[gq][co][t][q];
    };
    auto add = [&](int slot) {
#pragma unroll
      for (int co = 0; co < 3; co++)
#pragma unroll
        for (int t = 0; t < NT16; t++)
#pragma unroll
          for (int q = 0; q < 4; q++) acc[gq][co][t][q] += red[slot * SL + ((co * NT16 + t) * 4 + q) * 64 + lane];
    };
    if (gq > 0) __syncthreads();
#pragma unroll
    for (int h = NW / 2; h >= 1; h /= 2) {
      if (wave >= h && wave < 2 * h) put(wave - h);
      __syncthreads();
      if (wave < h) add(wave);
      if (h > 1) __syncthreads();
    }
    if (wave == 0 && (NRHS > 1 || (part >> 1) == 0)) {
#pragma unroll
      for (int co = 0; co < 3; co++) {
        const R so = co == 0 ? sgx : (co == 1 ? sgy : R(1));
#pragma unroll
        for (int t = 0; t < NT16; t++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int row = co * NZ + 16 * t + mma_row<R>(kq, q);
            vo[2 * (((2 * gq + rl) * 4 + pnt) * NZ3 + row) + reim] = acc[gq][co][t][q] * so;
          }
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < NRHS * 4 * nz3; i += NTH) {
    const int rp = C3_DIVNZ3(i), e = i - rp * nz3, p = rp & 3, ci = C3_DIVNZ(e), z = e - ci * nz;
    if (C3_MSET(rp >> 2) && !two) continue;
    T2 o = vin[rp * NZ3 + ci * NZ + z];
    if (!VAL(p)) o = zero;
    C3_ROWP(rp >> 2, p)[C3_CS(rp >> 2, ci) * nz + z] = o;
  }
#undef C3_VQL
#undef C3_VQP
#undef C3_DIVNZ
#undef C3_DIVNZ3
#undef C3_LDF
#undef C3_LDG
#undef C3_LDH
#undef C3_LDI
#undef C3_LDJ
#undef C3_LDK
#undef C3_PREFETCH
#undef C3_STX
#undef C3_STF
#undef C3_STG
#undef C3_STH
#undef C3_STI
#undef C3_STJ
#undef C3_STK
#undef C3_SETS
#undef C3_STASH
#undef PX
#undef PY
#undef VAL
#undef C3_MSET
#undef C3_ROWP
#undef C3_CS
#undef C3_LIVE
}
// Minimum waves per SIMD asked of the compiler for the unpadded kernel: four
// for complex64 at NZ = 48 (one or two right-hand sides fit in 128 registers
// a lane without spilling, against three waves otherwise); no request for
// the others, whose register use the plain launch bounds set (a request for
// more waves made the FP64 NZ = 32 kernel, the complex64 NZ = 32 and 64
// kernels and the padded complex64 NZ = 48 kernels spill)
template <typename R, int NZ, int NRHS> constexpr int c3_waves() {
  return (sizeof(R) == 4 && NZ == 48 && NRHS <= 2) ? 4 : 1;
}
// Nz a multiple of 16 (NZ = Nz)
template <typename R, int NZ, int NRHS, int DEEP, int NW, bool MIR = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(c3_waves<R, NZ, NRHS>()))) void k_contract3(cx_t<R>* __restrict__ S, int64_t sst,
                                                   const cx_t<R>* __restrict__ G, Geo g, int64_t BS) {
  contract3_body<R, NZ, NRHS, DEEP, NW, false, MIR>(S, sst, G, g, BS);
}
// other Nz from 8 to 64: the tiles of the next multiple of 16, zero-padded
// (at least two waves per SIMD: the padded paired kernels for FP64 at NZ = 48
// and complex64 at NZ = 64 otherwise take more than 256 registers a lane)
template <typename R, int NZ, int NRHS, int DEEP, int NW, bool MIR = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2))) void k_contract3_pad(cx_t<R>* __restrict__ S, int64_t sst,
                                                       const cx_t<R>* __restrict__ G, Geo g, int64_t BS) {
  contract3_body<R, NZ, NRHS, DEEP, NW, true, MIR>(S, sst, G, g, BS);
}

// ------------------------------------------------------------ dispatch ---
// The layer-split matrix-core contraction of NRHS spectra (rhs r at S + r*sst)
// through one tensor read: one or two for Nz <= 64, four or eight for
// Nz <= 32 (register budget).  Nz that is not a multiple of 16 runs the next
// multiple's kernel with zero-padded tiles.  False for other cases.
// MIR (NRHS = 2): one vector and its x-y mirror (op_mirror_ok), blocks over kx <= ky.
template <typename R, int NRHS, bool MIR = false>
static bool launch_contract3(const Operator& op, const cx_t<R>* G, cx_t<R>* S, int64_t sst, const Geo& g,
                             hipStream_t s) {
  const unsigned nblk = MIR ? (unsigned)(op.Nx * (op.Nx + 1) / 2) : (unsigned)(op.Nx * op.Ny);
  auto go = [&](auto kern, size_t sm, int nw) {
    G2G_CHECK(sm <= 160 * 1024, -3, "matrix-core contraction does not fit in LDS for this Nz");
    if (sm > 65536)
      G2G_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(64 * nw), sm, s, S, sst, G, g, op.BS);
  };
  // measured choices: FP64 keeps one component in flight (Nz = 32: four
  // waves per SIMD instead of three, 1-3% faster) and at Nz = 64 uses eight
  // waves (9-11% faster); complex64 at Nz = 64 stays at four waves (eight
  // were 12-18% slower)
  // (the point operand read from LDS per step instead of registers, and two
  // FP64 components in flight at four waves per SIMD, measured within 1%)
  constexpr bool F64 = sizeof(R) == 8;
  constexpr int DP = F64 ? 1 : 2, W64 = F64 ? 8 : 4;
  if (!c3_supported(op.Nz, NRHS)) return false;
  const int nzp = (op.Nz + 15) / 16 * 16;
  const bool pad = nzp != op.Nz;
  auto run = [&](auto ZC, auto DC, auto WC) {
    constexpr int Z = decltype(ZC)::value, D = decltype(DC)::value, W = decltype(WC)::value;
    if (pad) go(k_contract3_pad<R, Z, NRHS, D, W, MIR>, C3Cfg<R, Z, NRHS, W>::smem, W);
    else go(k_contract3<R, Z, NRHS, D, W, MIR>, C3Cfg<R, Z, NRHS, W>::smem, W);
  };
  using std::integral_constant;
  switch (nzp) {
    case 16: run(integral_constant<int, 16>{}, integral_constant<int, 2>{}, integral_constant<int, 4>{}); break;
    case 32: run(integral_constant<int, 32>{}, integral_constant<int, DP>{}, integral_constant<int, 4>{}); break;
    default:
      if constexpr (NRHS <= 2) {
        if (nzp == 48) run(integral_constant<int, 48>{}, integral_constant<int, 2>{}, integral_constant<int, 4>{});
        else run(integral_constant<int, 64>{}, integral_constant<int, DP>{}, integral_constant<int, W64>{});
      }
      break;
  }
  return true;
}
// ------------------------------------------- one vector, small-block MFMA ---
// The product of one vector with every output column a point.  The 16x16x4
// tiles above fill eight of their sixteen operand columns when one vector is
// contracted (four points x Re/Im), so half their matrix-core work repeats.
// Here every block of v_mfma_f32_4x4x1_16b (sixteen 4x4 blocks, each with its
// own operands) and of v_mfma_f64_4x4x4_4b (four 4x4x4 blocks) has the four
// points as its columns and one (FP32) or four (FP64) source layers as its
// depth, so no column repeats:
//   Re out += Gr Re v - Gi Im v,  Im out += Gr Im v + Gi Re v
// (four real products per tensor element).  Lane layouts (measured):
//   f32 4x4x1_16b: block b = lane/4; lane 4b+i supplies A_b[row i] and
//     B_b[col i]; lane 4b+c receives D_b[row q][col c] in register q.
//   f64 4x4x4_4b: lane 16k+4b+m supplies A_b[row m][k], B_b[k][col m] and
//     receives D_b[row k][col m].
// The whole tensor block of the (kx,ky) point (six components) is copied into
// LDS by direct global-to-LDS loads issued at once, so a block waits for one
// memory latency instead of one per component: the packed symmetric
// components as stored (16-byte lanes), the full xz and yz components
// gathered into an XOR-swizzled column layout, element (row, col) at
// col*NZ + (row ^ (col & SWM)), whose columns (A v) and rows (A^T v) are both
// read without bank conflicts.  Waves split the source layers; partial sums
// are added through LDS in a fixed order (results do not depend on
// scheduling).
template <typename R, int NZ, int NW_, bool MIR> struct C1Cfg {
  static constexpr bool F32 = sizeof(R) == 4;
  static constexpr int NW = NW_, NTH = 64 * NW, JW = NZ / NW;
  static constexpr int NSET = MIR ? 2 : 1;  // point sets per block: the quad and its mirror
  static constexpr int P = NZ * (NZ + 1) / 2;
  static constexpr int ES = (int)sizeof(cx_t<R>);
  static constexpr int SWM = (F32 && NZ >= 32) ? 31 : 15;
  // FP32: one instruction = one output component, G source layers x NZ rows
  static constexpr int G = 64 / NZ;
  // FP64: one instruction = sixteen output rows (four blocks of four)
  static constexpr int NI = F32 ? 3 : 3 * NZ / 16;
  static constexpr int VS = 3 * NZ * 4;                           // point-row values per set
  static constexpr int VIN_B = NSET * VS * ES;                    // [set][ci][z][point]
  static constexpr int SYM_B = (4 * P * ES + 1023) / 1024 * 1024;  // whole 1 KB copies
  static constexpr int FULL_B = 2 * NZ * NZ * ES;
  // partial sums per wave: FP32 after the G layer offsets are added across
  // lanes (4 NZ values per output component and part), FP64 64 per tile row set
  static constexpr int RED_B = NSET * NW * NI * 2 * (F32 ? 4 * NZ : 64) * (int)sizeof(R);
  static constexpr int smem = VIN_B + (SYM_B + FULL_B > RED_B ? SYM_B + FULL_B : RED_B);
};

#define G2G_LDS(p) ((__attribute__((address_space(3))) void*)(p))
// cache policy of the tensor copies: non-temporal (each apply reads the tensor
// once; measured 6% faster for complex64 at 256x256x32, 8% for FP64 at
// 128x128x32, than the default policy)
constexpr int C1_NT = 2;
// f(integral_constant<int, i>) for i = 0..N-1
template <typename F, int... Is> __device__ __forceinline__ void static_for_seq(F& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F& f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// MIR: the x-y mirror of a square lateral grid with dx = dy.  Swapping the x
// and y axes maps the layered-earth tensor onto itself, so the block of
// (ky, kx) is the block of (kx, ky) with x and y exchanged:
//   G0(ky, kx) = P G0(kx, ky) P,  P = the x <-> y component swap.
// A block then serves the quad (kx, ky) and the mirrored quad (ky, kx)
// (kx < ky; blocks cover kx <= ky) from one tensor read.  With
// P D_p P = D_q (q: p with its two sign bits exchanged), point q of the
// mirrored set is the point whose row sits at the transposed coordinates
// (PY(q), PX(q)); its input is read x <-> y exchanged and its output stored
// the same way, and both take the signs of point q:
//   out'(q) = P D_q G0 D_q P v'(q).
// waves per SIMD asked of the compiler: four for the mirrored FP32 kernels,
// whose 39 KB of LDS at Nz = 32 fit four blocks per CU (their registers
// otherwise allowed three)
template <typename R, bool MIR> constexpr int c1_waves() { return (sizeof(R) == 4 && MIR) ? 4 : 1; }
template <typename R, int NZ, bool PAD, int NW_, bool MIR>
__global__ __launch_bounds__(64 * NW_) __attribute__((amdgpu_waves_per_eu(c1_waves<R, MIR>()))) void k_contract1(cx_t<R>* __restrict__ S, const cx_t<R>* __restrict__ G, Geo g,
                                                        int64_t BS) {
  using T2 = cx_t<R>;
  using C = C1Cfg<R, NZ, NW_, MIR>;
  constexpr bool F32 = C::F32;
  constexpr int NW = C::NW, NTH = C::NTH, JW = C::JW, ES = C::ES, SWM = C::SWM, NI = C::NI, NSET = C::NSET;
  constexpr int VS = C::VS;
  static_assert(NZ == 16 || NZ == 32, "16 or 32 layers (fewer: zero-padded)");
  static_assert(JW % (F32 ? C::G : 4) == 0, "whole source-layer groups per wave");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T2* vin = reinterpret_cast<T2*>(smem_raw);
  const T2* sym = reinterpret_cast<const T2*>(smem_raw + C::VIN_B);
  const T2* full = reinterpret_cast<const T2*>(smem_raw + C::VIN_B + C::SYM_B);
  R* red = reinterpret_cast<R*>(smem_raw + C::VIN_B);  // after the products, over the tensor
  const int nz = PAD ? g.Nz : NZ, nz3 = 3 * nz;
  const int Pn = PAD ? nz * (nz + 1) / 2 : C::P;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int kx, ky;
  if constexpr (MIR) {
    // block t -> (kx, ky) with kx <= ky, t = ky (ky + 1) / 2 + kx
    const int t = blockIdx.x;
    int y = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((y + 1) * (y + 2) / 2 <= t) y++;
    while (y * (y + 1) / 2 > t) y--;
    ky = y;
    kx = t - y * (y + 1) / 2;
  } else {
    // blocks walk ky fastest: concurrent blocks touch neighbouring spectrum rows
    kx = blockIdx.x / g.Ny;
    ky = blockIdx.x % g.Ny;
  }
  const int b = kx + g.Nx * ky;
  const bool two = MIR && kx != ky;  // the diagonal maps onto itself: one set
  const int Nx2 = 2 * g.Nx, Ny2 = 2 * g.Ny, gNx = g.Nx, gNy = g.Ny;
#define PX(p) (((p) & 1) ? (kx == 0 ? gNx : Nx2 - kx) : kx)
#define PY(p) (((p) & 2) ? (ky == 0 ? gNy : Ny2 - ky) : ky)
#define VAL(p) ((!((p) & 1) || kx > 0) && (!((p) & 2) || ky > 0))
  // spectrum row of point q of set st (the mirrored set at transposed coordinates)
#define ROW(st, q) ((st) == 0 ? (int64_t)PX(q) * Ny2 + PY(q) : (int64_t)PY(q) * Ny2 + PX(q))
  const T2* blk = G + (int64_t)b * BS;
  const T2 zero = mk<T2>(0, 0);
  // 1. the point rows into registers (issued first: waiting for them does
  //    not wait for the tensor copies issued after them).  Wave w takes the
  //    rows w*RPW .. w*RPW + RPW - 1 (row = 4 set + point), so a row's base
  //    and validity are wave-uniform; lane l takes elements l and l + 64 of
  //    its 3 NZ (component-major)
  constexpr int NROW = NSET * 4, RPW = NROW / NW, EPL = (3 * NZ + 63) / 64;
  static_assert(NROW % NW == 0, "whole point rows per wave");
  T2 vq[RPW][EPL];
#pragma unroll
  for (int rr = 0; rr < RPW; rr++) {
    const int rho = wave * RPW + rr, st = rho >> 2, q = rho & 3;
    const bool live = (st == 0 || two) && VAL(q);
    const T2* row = S + ROW(st, q) * nz3;
#pragma unroll
    for (int k = 0; k < EPL; k++) {
      const int e = lane + 64 * k, ci = e / NZ, z = e - ci * NZ;
      const int cs = (st == 0 || ci == 2) ? ci : 1 - ci;  // mirrored set: x <-> y
      vq[rr][k] = zero;
      if (live && e < 3 * NZ && z < nz) vq[rr][k] = row[cs * nz + z];
    }
  }
  // 2. the tensor block into LDS
  {
    const unsigned char* gb = reinterpret_cast<const unsigned char*>(blk);
    const unsigned ngr = 4 * Pn * ES / 16;  // 16-byte pieces of the packed components (4 Pn is even)
    const unsigned nts = (ngr + 63) / 64;
    for (unsigned t = wave; t < nts; t += NW) {
      const unsigned gi = min(t * 64 + lane, ngr - 1);
      __builtin_amdgcn_global_load_lds((const void*)(gb + 16 * gi), G2G_LDS(smem_raw + C::VIN_B + t * 1024), 16, 0,
                                       C1_NT);
    }
    // full components, swizzled: copy t fills positions t PER .. t PER + PER - 1
    // (component-major, then column-major); unsigned offsets from the
    // component base keep the per-lane work to a few operations
    const unsigned char* fb = reinterpret_cast<const unsigned char*>(blk + 4 * Pn);
    constexpr unsigned PER = F32 ? 32 : 64;  // complex values per copy instruction
    constexpr unsigned NTF = 2 * NZ * NZ / PER;
    const unsigned lp = F32 ? (unsigned)lane >> 1 : (unsigned)lane;
    const unsigned part = F32 ? 4u * (lane & 1) : 0u;
    for (unsigned t = wave; t < NTF; t += NW) {
      const unsigned pos = t * PER + lp;
      const unsigned comp = pos / (NZ * NZ), loc = pos % (NZ * NZ);
      const unsigned col = loc / NZ, row = (loc % NZ) ^ (col & SWM);
      // padded rows and columns copy a real element (finite; its products
      // meet zero operands or land in rows that are not stored)
      const unsigned rc = PAD ? min(row, (unsigned)nz - 1) : row, cc = PAD ? min(col, (unsigned)nz - 1) : col;
      const unsigned off = ((comp * nz + cc) * nz + rc) * ES + part;
      unsigned char* dst = smem_raw + C::VIN_B + C::SYM_B + t * (PER * ES);
      if constexpr (F32)
        __builtin_amdgcn_global_load_lds((const void*)(fb + off), G2G_LDS(dst), 4, 0, C1_NT);
      else
        __builtin_amdgcn_global_load_lds((const void*)(fb + off), G2G_LDS(dst), 16, 0, C1_NT);
    }
  }
  // 3. point rows with the point signs (x times -1 for -kx, y for -ky; 0/1
  //    flags, no branch), layer-major with the four points innermost
#pragma unroll
  for (int rr = 0; rr < RPW; rr++) {
    const int rho = wave * RPW + rr, st = rho >> 2, q = rho & 3;
#pragma unroll
    for (int k = 0; k < EPL; k++) {
      const int e = lane + 64 * k, ci = e / NZ, z = e - ci * NZ;
      const int flip = ((ci == 0) & (q & 1)) | ((ci == 1) & ((q >> 1) & 1));
      const R sg = R(1 - 2 * flip);
      if (e < 3 * NZ) vin[st * VS + (ci * NZ + z) * 4 + q] = mk<T2>(sg * vq[rr][k].x, sg * vq[rr][k].y);
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  // 4. products.  Tensor element (r, j) of output component co and input ci:
  //    ci = x: Sxx, Sxy, -Axz^T;  ci = y: Sxy, Syy, -Ayz^T;  ci = z: Axz, Ayz, Szz
  //    A lane's offsets of element (r, j) in a packed component (ps), in a
  //    full one read down a column (pd) and along a row (pt) do not depend
  //    on the component: computed once per source layer, the fetches then
  //    add compile-time component offsets
  struct Off { int ps, pd, pt; bool in; };
  auto offsets = [&](int r, int j) -> Off {
    Off o;
    o.ps = j <= r ? j + r * (r + 1) / 2 : r + j * (j + 1) / 2;
    o.pd = j * NZ + (r ^ (j & SWM));
    o.pt = r * NZ + (j ^ (r & SWM));
    o.in = !PAD || j < nz;
    return o;
  };
  auto fetch = [&](auto CI, auto CO, const Off& o) -> T2 {
    constexpr int ci = decltype(CI)::value, co = decltype(CO)::value;
    if constexpr (ci == 2 && co < 2) {
      return full[co * NZ * NZ + o.pd];
    } else if constexpr (ci < 2 && co == 2) {
      const T2 a = full[ci * NZ * NZ + o.pt];
      return mk<T2>(-a.x, -a.y);
    } else {
      constexpr int comp = ci == 2 ? 3 : (ci != co ? 1 : (ci == 0 ? 0 : 2));
      T2 a = sym[comp * Pn + o.ps];
      if (PAD && !o.in) a = zero;  // beyond the packed rows: not stored
      return a;
    }
  };
  using std::integral_constant;
  if constexpr (F32) {
    using V4 = v4t<float>::t;
    V4 are[NSET][3], aim[NSET][3];
#pragma unroll
    for (int st = 0; st < NSET; st++)
#pragma unroll
      for (int co = 0; co < 3; co++) are[st][co] = aim[st][co] = V4{0, 0, 0, 0};
    // lane: A row r, source-layer offset gl, point column pc
    const int r = lane % NZ, gl = lane / NZ, pc = lane & 3;
    constexpr int NJG = JW / C::G;
    Off off[NJG];
#pragma unroll
    for (int jg = 0; jg < NJG; jg++) off[jg] = offsets(r, wave * JW + jg * C::G + gl);
    auto ci_pass = [&](auto CI) {
      constexpr int ci = decltype(CI)::value;
#pragma unroll
      for (int jg = 0; jg < NJG; jg++) {
        const int j = wave * JW + jg * C::G + gl;
        const T2 a0 = fetch(CI, integral_constant<int, 0>{}, off[jg]);
        const T2 a1 = fetch(CI, integral_constant<int, 1>{}, off[jg]);
        const T2 a2 = fetch(CI, integral_constant<int, 2>{}, off[jg]);
#pragma unroll
        for (int st = 0; st < NSET; st++) {
          const T2 bv = vin[st * VS + (ci * NZ + j) * 4 + pc];
          const float nbi = -bv.y;
          are[st][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(a0.x, bv.x, are[st][0], 0, 0, 0);
          are[st][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(a1.x, bv.x, are[st][1], 0, 0, 0);
          are[st][2] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.x, bv.x, are[st][2], 0, 0, 0);
          aim[st][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(a0.x, bv.y, aim[st][0], 0, 0, 0);
          aim[st][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(a1.x, bv.y, aim[st][1], 0, 0, 0);
          aim[st][2] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.x, bv.y, aim[st][2], 0, 0, 0);
          are[st][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(a0.y, nbi, are[st][0], 0, 0, 0);
          are[st][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(a1.y, nbi, are[st][1], 0, 0, 0);
          are[st][2] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.y, nbi, are[st][2], 0, 0, 0);
          aim[st][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(a0.y, bv.x, aim[st][0], 0, 0, 0);
          aim[st][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(a1.y, bv.x, aim[st][1], 0, 0, 0);
          aim[st][2] = __builtin_amdgcn_mfma_f32_4x4x1f32(a2.y, bv.x, aim[st][2], 0, 0, 0);
        }
      }
    };
    ci_pass(integral_constant<int, 0>{});
    ci_pass(integral_constant<int, 1>{});
    ci_pass(integral_constant<int, 2>{});
    // the G layer offsets of a row block sit in lanes NZ apart: add them
    // across lanes (offsets 0+1, then +2+3 at NZ = 16), lanes below NZ keep
    // the sums
#pragma unroll
    for (int st = 0; st < NSET; st++)
#pragma unroll
      for (int co = 0; co < 3; co++)
#pragma unroll
        for (int o = NZ; o < 64; o <<= 1)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            are[st][co][q] += __shfl_xor(are[st][co][q], o);
            aim[st][co][q] += __shfl_xor(aim[st][co][q], o);
          }
    __syncthreads();
    if (lane < NZ) {
#pragma unroll
      for (int st = 0; st < NSET; st++)
#pragma unroll
        for (int co = 0; co < 3; co++) {
          R* dst = red + (((st * NW + wave) * 3 + co) * 2) * (4 * NZ) + lane * 4;
          *reinterpret_cast<V4*>(dst) = are[st][co];
          *reinterpret_cast<V4*>(dst + 4 * NZ) = aim[st][co];
        }
    }
  } else {
    double are[NSET][NI], aim[NSET][NI];
#pragma unroll
    for (int st = 0; st < NSET; st++)
#pragma unroll
      for (int I = 0; I < NI; I++) are[st][I] = aim[st][I] = 0.0;
    // lane 16k + 4b + m: source layer k of the quad, block b, A row / B column m
    const int kq = lane >> 4, bq = (lane >> 2) & 3, m = lane & 3;
    // element offsets per (source-layer quad, row set): the row set of
    // instruction I depends on I only through (4 I) mod (NZ / 4)
    constexpr int NJQ = JW / 4, NRS = (NZ / 4) / 4;
    Off off[NJQ][NRS];
#pragma unroll
    for (int jq = 0; jq < NJQ; jq++)
#pragma unroll
      for (int rs = 0; rs < NRS; rs++) off[jq][rs] = offsets(4 * (4 * rs + bq) + m, wave * JW + 4 * jq + kq);
    auto ci_pass = [&](auto CI) {
      constexpr int ci = decltype(CI)::value;
#pragma unroll
      for (int jq = 0; jq < NJQ; jq++) {
        const int j = wave * JW + 4 * jq + kq;
        T2 bv[NSET];
#pragma unroll
        for (int st = 0; st < NSET; st++) bv[st] = vin[st * VS + (ci * NZ + j) * 4 + m];
        auto rows = [&](auto II) {
          constexpr int I = decltype(II)::value, co = (4 * I) / (NZ / 4), rs = I % NRS;
          const T2 a = fetch(CI, integral_constant<int, co>{}, off[jq][rs]);
#pragma unroll
          for (int st = 0; st < NSET; st++) {
            are[st][I] = __builtin_amdgcn_mfma_f64_4x4x4f64(a.x, bv[st].x, are[st][I], 0, 0, 0);
            aim[st][I] = __builtin_amdgcn_mfma_f64_4x4x4f64(a.x, bv[st].y, aim[st][I], 0, 0, 0);
            are[st][I] = __builtin_amdgcn_mfma_f64_4x4x4f64(a.y, -bv[st].y, are[st][I], 0, 0, 0);
            aim[st][I] = __builtin_amdgcn_mfma_f64_4x4x4f64(a.y, bv[st].x, aim[st][I], 0, 0, 0);
          }
        };
        static_for<NI>(rows);
      }
    };
    ci_pass(integral_constant<int, 0>{});
    ci_pass(integral_constant<int, 1>{});
    ci_pass(integral_constant<int, 2>{});
    __syncthreads();
#pragma unroll
    for (int st = 0; st < NSET; st++)
#pragma unroll
      for (int I = 0; I < NI; I++) {
        red[(((st * NW + wave) * NI + I) * 2 + 0) * 64 + lane] = are[st][I];
        red[(((st * NW + wave) * NI + I) * 2 + 1) * 64 + lane] = aim[st][I];
      }
  }
  __syncthreads();
  // 5. sums over waves in a fixed order, output signs, store: the rows as
  //    in step 1 (wave-uniform row, lane-strided elements)
#pragma unroll
  for (int rr = 0; rr < RPW; rr++) {
    const int rho = wave * RPW + rr, st = rho >> 2, q = rho & 3;
    if (st == 1 && !two) continue;
    const bool live = VAL(q);
    T2* row = S + ROW(st, q) * nz3;
#pragma unroll
    for (int k = 0; k < EPL; k++) {
      const int e = lane + 64 * k, co = e / NZ, z = e - co * NZ;
      if (e >= 3 * NZ || z >= nz) continue;
      R sr = 0, si = 0;
      if constexpr (F32) {
#pragma unroll
        for (int w = 0; w < NW; w++) {
          const int at = (((st * NW + w) * 3 + co) * 2) * (4 * NZ) + (4 * (z >> 2) + q) * 4 + (z & 3);
          sr += red[at];
          si += red[at + 4 * NZ];
        }
      } else {
        const int rb = co * (NZ / 4) + (z >> 2), I = rb >> 2, bq = rb & 3;
#pragma unroll
        for (int w = 0; w < NW; w++) {
          const int at = (((st * NW + w) * NI + I) * 2) * 64 + 16 * (z & 3) + 4 * bq + q;
          sr += red[at];
          si += red[at + 64];
        }
      }
      // output sign as arithmetic on 0/1 flags: the short-circuit form
      // (co == 0 && p & 1) || (co == 1 && p & 2) selecting -sum compiled, for
      // FP32 with this toolchain, to code that kept +sum for co = 1, p & 2
      const int flip = ((co == 0) & (q & 1)) | ((co == 1) & ((q >> 1) & 1));
      const R so = R(1 - 2 * flip);
      const int cs = (st == 0 || co == 2) ? co : 1 - co;  // mirrored set: x <-> y
      row[cs * nz + z] = live ? mk<T2>(so * sr, so * si) : zero;
    }
  }
#undef PX
#undef PY
#undef VAL
#undef ROW
}

// waves per block at 32 layers (eight measured 1% slower in FP64 and 6%
// slower in complex64 at 256x256x32)
template <typename R> constexpr int C1_W32 = 4;
// Automatic use of the small-block kernel for one vector (contraction alone
// at 128x128, Nz = 8 to 32): FP64 at every Nz (3-30% faster than the 16x16x4
// tiles, except 1% slower at Nz = 28); complex64 from Nz = 24 (3-9% faster;
// below 17 layers the tensor stays in the 256 MB last-level cache and the
// 16x16x4 tiles were 10-20% faster, and at Nz = 17-23 the vector-ALU kernel
// stays faster)
template <typename R> bool c1_auto(int nz) { return nz >= (sizeof(R) == 8 ? 8 : 24) && nz <= 32; }
// The small-block kernel for Nz from 8 to 32 (padded to 16 or 32 layers),
// with the x-y mirror when mir is set (square grid, dx = dy, the library's
// own tensor).  False for other Nz.
template <typename R>
static bool launch_contract1(const Operator& op, const cx_t<R>* G, cx_t<R>* S, const Geo& g, bool mir,
                             hipStream_t s) {
  if (op.Nz < C3_MIN_NZ || op.Nz > 32) return false;
  const unsigned nblk = mir ? (unsigned)(op.Nx * (op.Nx + 1) / 2) : (unsigned)(op.Nx * op.Ny);
  auto go = [&](auto kern, int sm, int nw) {
    if (sm > 65536)
      G2G_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(64 * nw), sm, s, S, G, g, op.BS);
  };
  constexpr int W32 = C1_W32<R>;
  auto run = [&](auto MC) {
    constexpr bool M = decltype(MC)::value;
    if (op.Nz == 16) go(k_contract1<R, 16, false, 4, M>, C1Cfg<R, 16, 4, M>::smem, 4);
    else if (op.Nz == 32) go(k_contract1<R, 32, false, W32, M>, C1Cfg<R, 32, W32, M>::smem, W32);
    else if (op.Nz < 16) go(k_contract1<R, 16, true, 4, M>, C1Cfg<R, 16, 4, M>::smem, 4);
    else go(k_contract1<R, 32, true, W32, M>, C1Cfg<R, 32, W32, M>::smem, W32);
  };
  if (mir) run(std::true_type{});
  else run(std::false_type{});
  return true;
}
// Automatic kernel for one vector without the x-y mirror (with it, measured
// at 128x128 for Nz = 8 to 64: the small-block kernel where c1_auto says so,
// else the mirrored 16x16x4 tiles, which beat the vector-ALU kernel at every
// Nz, 1.5-3.4x): the matrix cores, except complex64 at
// padded layer counts where the vector-ALU kernel measured faster (128x128,
// contraction alone: Nz = 20 by 19%, Nz = 50 and 56 by 5-8%; the padded tiles
// won at Nz = 8, 12, 28, 40, tied at 24, and won in FP64 at every Nz
// measured), i.e. where padding adds more than 80% of the tile work or pads
// to 64 layers.
template <typename R> bool c3_auto(int nz) {
  const int nzp = (nz + 15) / 16 * 16;
  if (sizeof(R) == 8 || nzp == nz || nzp == 16) return true;
  return nzp < 64 && 5 * nzp * nzp <= 9 * nz * nz;
}
template <typename R, template <typename, int> class KF> struct Dummy {};

static Geo make_geo(const Operator& op, int nrx) {
  Geo g;
  g.Nx = op.Nx;
  g.Ny = op.Ny;
  g.Nz = op.Nz;
  g.Nz3 = 3 * op.Nz;
  g.nrx_log = ilog2(nrx);
  g.nc = op.ncell();
  for (int c = 0; c < 3; c++) g.slot[c] = op.comp_slot[c];
  g.px = op.px;
  g.py = op.py;
  g.chz = (op.Nz % nrx == 0) ? op.Nz / nrx : 0;
  return g;
}

template <typename R> static void make_twiddles(int L, cx_t<R>** dptr) {
  std::vector<cx_t<R>> h(L);
  for (int t = 0; t < L; t++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * t / L;
    h[t] = mk<cx_t<R>>((R)cosl(a), (R)sinl(a));
  }
  G2G_HIP(hipMalloc(dptr, sizeof(cx_t<R>) * L));
  G2G_HIP(hipMemcpy(*dptr, h.data(), sizeof(cx_t<R>) * L, hipMemcpyHostToDevice));
}

template <typename R> void op_alloc_buffers(Operator& op, OpBuffers<R>& b, bool with_tensor) {
  using T2 = cx_t<R>;
  const int Nx2 = 2 * op.Nx, Ny2 = 2 * op.Ny, Nz3 = 3 * op.Nz;
  b.half = use_half_fft<R>(Nx2, Ny2);
  b.NR = b.half ? nr_half_rt<R>(Nx2) : nr_rt<R>(Nx2);
  b.nchunk = (Nz3 + b.NR - 1) / b.NR;
  const size_t nX = (size_t)b.nchunk * b.NR * op.Ny * Nx2;
  const size_t nS = (size_t)Nx2 * Ny2 * Nz3;
  const size_t nc = op.ncell();
  b.bytes = 0;
  auto A = [&](auto** p, size_t n, size_t es) {
    G2G_HIP(hipMalloc((void**)p, n * es));
    b.bytes += n * es;
  };
  if (with_tensor) A(&b.G, (size_t)op.Nx * op.Ny * op.BS, sizeof(T2));
  A(&b.X, nX, sizeof(T2));
  A(&b.S, nS, sizeof(T2));
  A(&b.gsig, nc, sizeof(T2));
  A(&b.asiga, nc, sizeof(T2));
  A(&b.sq, (size_t)op.Nz, sizeof(R));
  A(&b.dzn, (size_t)op.Nz, sizeof(R));
  make_twiddles<R>(Nx2, &b.twx);
  make_twiddles<R>(Ny2, &b.twy);
  std::vector<R> dzn(op.Nz);
  for (int z = 0; z < op.Nz; z++) dzn[z] = (R)(1.0 / (op.dz[z] * 4.0 * op.Nx * op.Ny));
  G2G_HIP(hipMemcpy(b.dzn, dzn.data(), sizeof(R) * op.Nz, hipMemcpyHostToDevice));
}

template <typename R> void op_free_buffers(OpBuffers<R>& b) {
  auto F = [](auto*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
  };
  F(b.G); F(b.X); F(b.S); F(b.SB); F(b.gsig); F(b.asiga); F(b.sq); F(b.dzn); F(b.twx); F(b.twy);
}

// The forward (pre-scale, pad, x then y) and inverse (y then x, post-scale)
// transforms of one vector between its cells and the spectrum S, with the
// half-length kernels where b.half is set.
template <typename R>
static void fft_forward(const Operator& op, OpBuffers<R>& b, const cx_t<R>* vin, cx_t<R>* S, const Geo& g,
                        hipStream_t s) {
  const int Nx2 = 2 * op.Nx, Ny2 = 2 * op.Ny, Nz3 = 3 * op.Nz;
  const int nry = b.half ? nr_half_rt<R>(Ny2) : nr_rt<R>(Ny2);
  const int chy = (Nz3 + nry - 1) / nry;
  len_dispatch(Nx2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    constexpr bool RT = decltype(LC)::rt;
    if (!RT && b.half)
      hipLaunchKernelGGL((k_fwd_x_h<R, LL>), dim3(op.Ny, b.nchunk), dim3(NT), 0, s, vin, b.gsig, b.X, b.twx, g);
    else
      hipLaunchKernelGGL((k_fwd_x<R, LL, RT>), dim3(op.Ny, b.nchunk), dim3(NT), 0, s, vin, b.gsig, b.X, b.twx,
                         g);
  });
  len_dispatch(Ny2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    constexpr bool RT = decltype(LC)::rt;
    if (!RT && b.half)
      hipLaunchKernelGGL((k_fwd_y_h<R, LL>), dim3(Nx2, chy), dim3(NT), 0, s, b.X, S, b.twy, g);
    else
      hipLaunchKernelGGL((k_fwd_y<R, LL, false, RT>), dim3(Nx2, chy), dim3(NT), 0, s, b.X, S, b.twy, g);
  });
}

template <typename R>
static void fft_inverse(const Operator& op, OpBuffers<R>& b, const cx_t<R>* S, const cx_t<R>* vin,
                        cx_t<R>* vout, const Geo& g, hipStream_t s) {
  const int Nx2 = 2 * op.Nx, Ny2 = 2 * op.Ny, Nz3 = 3 * op.Nz;
  const int nry = b.half ? nr_half_rt<R>(Ny2) : nr_rt<R>(Ny2);
  const int chy = (Nz3 + nry - 1) / nry;
  len_dispatch(Ny2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    constexpr bool RT = decltype(LC)::rt;
    if (!RT && b.half)
      hipLaunchKernelGGL((k_inv_y_h<R, LL>), dim3(Nx2, chy), dim3(NT), 0, s, S, b.X, b.twy, g);
    else
      hipLaunchKernelGGL((k_inv_y<R, LL, false, RT>), dim3(Nx2, chy), dim3(NT), 0, s, S, b.X, b.twy, g);
  });
  len_dispatch(Nx2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    constexpr bool RT = decltype(LC)::rt;
    if (!RT && b.half)
      hipLaunchKernelGGL((k_inv_x_h<R, LL>), dim3(op.Ny, b.nchunk), dim3(NT), 0, s, b.X, vin, b.asiga, b.sq,
                         b.dzn, vout, b.twx, g);
    else
      hipLaunchKernelGGL((k_inv_x<R, LL, RT>), dim3(op.Ny, b.nchunk), dim3(NT), 0, s, b.X, vin, b.asiga, b.sq,
                         b.dzn, vout, b.twx, g);
  });
}

template <typename R>
void op_apply(Operator& op, OpBuffers<R>& b, const cx_t<R>* vin, cx_t<R>* vout, hipStream_t s,
              bool timed, bool mirror) {
  G2G_CHECK(op.have_tensor && op.have_sigma, -4, "operator not ready: tensor or conductivity missing");
  const int Nz3 = 3 * op.Nz;
  Geo g = make_geo(op, b.NR);
  // profiling: events around the three stages (flushed into op.cnt later)
  const bool prof = op.prof.on && !timed;
  const int prec = sizeof(R) == 8 ? 0 : 1;
  int pe[4] = {0, 0, 0, 0};
  if (prof) pe[0] = op.prof.record(s);
  if (timed) G2G_HIP(hipEventRecord(op.ev[0], s));
  fft_forward<R>(op, b, vin, b.S, g, s);
  if (timed) G2G_HIP(hipEventRecord(op.ev[1], s));
  if (prof) pe[1] = op.prof.record(s);
  if (op.Nz > 64) {
    // tiles of TJ source columns; the four points' rows and one tile in LDS
    int split_log = 0;
    while ((op.Nz << (split_log + 1)) <= NT) split_log++;
    const size_t es = sizeof(cx_t<R>), vb = es * 4 * (size_t)Nz3;
    const int TJ = (int)std::max<size_t>(4, std::min<size_t>(32, (96 * 1024 - vb) / (es * op.Nz)));
    const size_t sm = vb + es * (size_t)TJ * op.Nz;
    auto kern = k_contract_tiled<R>;
    if (sm > 65536)
      G2G_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    hipLaunchKernelGGL(kern, dim3(op.Nx * op.Ny), dim3(NT), sm, s, b.S, b.G, g, op.BS, split_log, TJ);
  } else {
    // vector ALU: rows = Nz layers, split lanes per row (power of two)
    int split_log = 0;
    while ((op.Nz << (split_log + 1)) <= NT) split_log++;
    const size_t sm = sizeof(cx_t<R>) * (4 * (size_t)Nz3 + 2 * (size_t)op.Nz * (op.Nz + 1));
    auto launch = [&](auto kern) {
      if (sm > 65536) {
        G2G_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024));
      }
      hipLaunchKernelGGL(kern, dim3(op.Nx * op.Ny), dim3(NT), sm, s, b.S, b.G, g, op.BS, split_log);
    };
    // matrix cores for Nz from 8 to 64 (complex64: where c3_auto says so)
    // unless a vector-ALU kernel is requested (FP64 at 128x128x32:
    // 0.28-0.30 ms against 0.33 ms)
    // the x-y mirror where it applies (modes 0, 5 and 6); small-block
    // kernel, then the mirrored 16x16x4 tiles, then the plain tiles
    const bool mir = mirror && op_mirror_ok(op);
    bool done = false;
    if (op.contraction == 4 || op.contraction == 5 || (op.contraction == 0 && c1_auto<R>(op.Nz)))
      done = launch_contract1<R>(op, b.G, b.S, g, op.contraction != 4 && mir, s);
    if (!done && mir && (op.contraction == 6 || op.contraction == 0) && c3_supported(op.Nz, 2))
      done = launch_contract3<R, 2, true>(op, b.G, b.S, 0, g, s);
    if (!done && (op.contraction == 2 || op.contraction == 6 || (op.contraction == 0 && c3_auto<R>(op.Nz))))
      done = launch_contract3<R, 1>(op, b.G, b.S, 0, g, s);
    if (done) {
    } else if (op.contraction == 3) {
      if (op.Nz * op.Nz <= 4 * NT) launch(k_contract<R, 4>);
      else launch(k_contract<R, 16>);
    } else if (op.Nz * op.Nz <= 4 * NT) {
      switch (op.Nz) {  // compile-time Nz for the common layer counts (one to four layers per lane)
        case 16: launch(k_contract2<R, 4, NT, 16>); break;
        case 32: launch(k_contract2<R, 4, NT, 32>); break;
        default: launch(k_contract2<R, 4, NT, 0>); break;
      }
    } else {
      // Nz > 32: 512 threads (8 lanes per receiver row at Nz = 64) keep twice
      // the waves per CU that the LDS budget (one or two blocks) allows
      int sl2 = 0;
      while ((op.Nz << (sl2 + 1)) <= 2 * NT) sl2++;
      // (a compile-time Nz = 64 unrolls eight source layers per lane and
      // spills registers: measured 40% slower in FP64, so Nz stays a runtime value)
      auto kern = k_contract2<R, 8, 2 * NT, 0>;
      if (sm > 65536) {
        G2G_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024));
      }
      hipLaunchKernelGGL(kern, dim3(op.Nx * op.Ny), dim3(2 * NT), sm, s, b.S, b.G, g, op.BS, sl2);
    }
  }
  if (timed) G2G_HIP(hipEventRecord(op.ev[2], s));
  if (prof) pe[2] = op.prof.record(s);
  fft_inverse<R>(op, b, b.S, vin, vout, g, s);
  if (timed) G2G_HIP(hipEventRecord(op.ev[3], s));
  if (prof) {
    pe[3] = op.prof.record(s);
    op.prof.mark(Prof::FWD, prec, 1, pe[0], pe[1]);
    op.prof.mark(Prof::CONTR, prec, 1, pe[1], pe[2]);
    op.prof.mark(Prof::BWD, prec, 1, pe[2], pe[3]);
  }
  G2G_HIP(hipGetLastError());
}

bool op_pair_supported(const Operator& op) { return c3_supported(op.Nz, 2); }
int op_batch_max(const Operator& op) {
  return c3_supported(op.Nz, 8) ? 8 : (c3_supported(op.Nz, 2) ? 2 : 1);
}

template <typename T2> __global__ void k_zero(T2* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) a[i] = mk<T2>(0, 0);
}

// k right-hand sides (1 <= k <= 8, k <= op_batch_max) through one tensor read:
// each vector is transformed into its own slot of the batch spectrum
// buffer, one contraction with k operand columns serves all of them, then
// each is transformed back.  k = 3, 5, 6, 7 run the next larger contraction
// on zero spectra in the unused slots.
template <typename R>
static void op_apply_k(Operator& op, OpBuffers<R>& b, int k, const cx_t<R>* const* vins,
                       cx_t<R>* const* vouts, hipStream_t s, bool mirror = false) {
  using T2 = cx_t<R>;
  const int Nx2 = 2 * op.Nx, Ny2 = 2 * op.Ny, Nz3 = 3 * op.Nz;
  const int kk = k <= 1 ? 1 : (k <= 2 ? 2 : (k <= 4 ? 4 : 8));
  const int64_t nS = (int64_t)Nx2 * Ny2 * Nz3;
  if (b.sb_cap < kk) {
    if (b.SB) {
      (void)hipFree(b.SB);
      b.bytes -= sizeof(T2) * (size_t)nS * b.sb_cap;
    }
    G2G_HIP(hipMalloc((void**)&b.SB, sizeof(T2) * (size_t)nS * kk));
    b.bytes += sizeof(T2) * (size_t)nS * kk;
    b.sb_cap = kk;
  }
  // unused slots hold zero spectra (a fill kernel, not a memset node, so
  // that a caller may capture the apply in a graph)
  if (kk > k) {
    const int64_t nz = nS * (kk - k);
    hipLaunchKernelGGL((k_zero<T2>), dim3((unsigned)std::min<int64_t>(4096, (nz + NT - 1) / NT)), dim3(NT), 0, s,
                       b.SB + nS * k, nz);
  }
  Geo g = make_geo(op, b.NR);
  const bool prof = op.prof.on;
  const int prec = sizeof(R) == 8 ? 0 : 1;
  int pe[4] = {0, 0, 0, 0};
  if (prof) pe[0] = op.prof.record(s);
  for (int r = 0; r < k; r++) fft_forward<R>(op, b, vins[r], b.SB + nS * r, g, s);
  if (prof) pe[1] = op.prof.record(s);
  bool ok = false;
  switch (kk) {
    case 1:
      // (no x-y mirror in the batched applies, as for their single vectors)
      ok = (op.contraction == 0 && c1_auto<R>(op.Nz)) ? launch_contract1<R>(op, b.G, b.SB, g, false, s)
                                                        : launch_contract3<R, 1>(op, b.G, b.SB, nS, g, s);
      break;
    case 2:
      // FP64: two mirrored one-vector contractions where the small-block
      // kernel serves one vector.  Each reads the tensor once per (kx,ky),
      // (ky,kx) pair of quads, together as much as the two-vector 16x16x4
      // kernel reads (256x256x32 contraction 2 x 599 against 1279 us; paired
      // FP64 solves 0.376 -> 0.372 s).  In complex64 the same change made the
      // paired mixed solves 2.5% slower, so they keep the two-vector kernel.
      if (sizeof(R) == 8 && mirror && op.contraction == 0 && op_mirror_ok(op) && c1_auto<R>(op.Nz))
        ok = launch_contract1<R>(op, b.G, b.SB, g, true, s) && launch_contract1<R>(op, b.G, b.SB + nS, g, true, s);
      else
        ok = launch_contract3<R, 2>(op, b.G, b.SB, nS, g, s);
      break;
    case 4: ok = launch_contract3<R, 4>(op, b.G, b.SB, nS, g, s); break;
    default:
      // eight per read in complex64 only (op_apply_batch caps FP64 at four)
      if constexpr (sizeof(R) == 4) ok = launch_contract3<R, 8>(op, b.G, b.SB, nS, g, s);
      break;
  }
  G2G_CHECK(ok, -3, "batched contraction not available for this Nz");
  if (prof) pe[2] = op.prof.record(s);
  for (int r = 0; r < k; r++) fft_inverse<R>(op, b, b.SB + nS * r, vins[r], vouts[r], g, s);
  if (prof) {
    pe[3] = op.prof.record(s);
    op.prof.mark(Prof::FWD, prec, k, pe[0], pe[1]);
    op.prof.mark(Prof::CONTR, prec, k, pe[1], pe[2]);
    op.prof.mark(Prof::BWD, prec, k, pe[2], pe[3]);
  }
  G2G_HIP(hipGetLastError());
}

// Two right-hand sides through one tensor read (op_pair_supported).
template <typename R>
void op_apply2(Operator& op, OpBuffers<R>& b, const cx_t<R>* vin0, const cx_t<R>* vin1, cx_t<R>* vout0,
               cx_t<R>* vout1, hipStream_t s) {
  G2G_CHECK(op.have_tensor && op.have_sigma, -4, "operator not ready: tensor or conductivity missing");
  G2G_CHECK(op_pair_supported(op), -3, "two-vector apply needs Nz from 8 to 64");
  const cx_t<R>* vins[2] = {vin0, vin1};
  cx_t<R>* vouts[2] = {vout0, vout1};
  op_apply_k<R>(op, b, 2, vins, vouts, s, true);
}
template void op_apply2<double>(Operator&, OpBuffers<double>&, const zd*, const zd*, zd*, zd*, hipStream_t);
template void op_apply2<float>(Operator&, OpBuffers<float>&, const zf*, const zf*, zf*, zf*, hipStream_t);

// nrhs vectors stored one after another (3 Nx Ny Nz complex each), in
// batches of at most op_batch_max(op) per tensor read.
template <typename R>
void op_apply_batch(Operator& op, OpBuffers<R>& b, int nrhs, const cx_t<R>* vin, cx_t<R>* vout,
                    hipStream_t s) {
  G2G_CHECK(op.have_tensor && op.have_sigma, -4, "operator not ready: tensor or conductivity missing");
  G2G_CHECK(nrhs >= 1, -3, "batched apply needs at least one vector");
  const int64_t nv = 3 * (int64_t)op.ncell();
  // FP64 takes at most four per tensor read: an eight-vector FP64 kernel
  // held 360 registers per lane (one wave per SIMD) and measured 22% slower
  // per vector than four at 128x128x32
  const int kmax = std::min(op_batch_max(op), sizeof(R) == 8 ? 4 : 8);
  for (int r0 = 0; r0 < nrhs;) {
    const int k = std::min(kmax, nrhs - r0);
    if (k == 1) {
      op_apply<R>(op, b, vin + nv * r0, vout + nv * r0, s, false, false);
    } else {
      const cx_t<R>* vins[8];
      cx_t<R>* vouts[8];
      for (int r = 0; r < k; r++) {
        vins[r] = vin + nv * (r0 + r);
        vouts[r] = vout + nv * (r0 + r);
      }
      op_apply_k<R>(op, b, k, vins, vouts, s);
    }
    r0 += k;
  }
}
template void op_apply_batch<double>(Operator&, OpBuffers<double>&, int, const zd*, zd*, hipStream_t);
template void op_apply_batch<float>(Operator&, OpBuffers<float>&, int, const zf*, zf*, hipStream_t);

// ------------------------------------------------ allocation-free batch ---
// Batched applies for callers that must not allocate device memory (code
// under graph capture, serving loops): the x-transformed rows X and the
// spectrum slots live in one file-scope device array, `scratch`, instead of
// the operator's buffers, which the first batched apply allocates.  Each GPU
// holds its own copy of the array from the moment the library's code is
// loaded on it until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is a stand-in, not a device address): the
// address of each device's copy is looked up once, when the first operator
// is created on that device, and kept in scratch_dev[device].
static __device__ __align__(256) zd scratch[NOALLOC_SCRATCH_BYTES / sizeof(zd)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_dev[MAX_DEVICES] = {};
}  // namespace

void op_noalloc_setup(int device) {
  G2G_CHECK(device >= 0 && device < MAX_DEVICES, -3, "device index beyond the library's table");
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_dev[device]) return;
  void* p = nullptr;
  G2G_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  G2G_CHECK(p != nullptr, -100, "no device address for the scratch array");
  scratch_dev[device] = p;
}

template <typename R> int op_noalloc_batch_max(const Operator& op, const OpBuffers<R>& b, size_t* need1) {
  const size_t es = sizeof(cx_t<R>);
  const size_t nX = (size_t)b.nchunk * b.NR * op.Ny * 2 * op.Nx;
  const size_t nS = (size_t)4 * op.Nx * op.Ny * 3 * op.Nz;
  if (need1) *need1 = es * (nX + nS);
  const int kmax = std::min(op_batch_max(op), sizeof(R) == 8 ? 4 : 8);
  int k = 0;
  for (int kk : {1, 2, 4, 8})
    if (kk <= kmax && es * (nX + nS * kk) <= NOALLOC_SCRATCH_BYTES) k = kk;
  return k;
}

template <typename R>
void op_apply_batch_noalloc(Operator& op, OpBuffers<R>& b, int nrhs, const cx_t<R>* vin, cx_t<R>* vout,
                            hipStream_t s) {
  using T2 = cx_t<R>;
  G2G_CHECK(nrhs >= 1, -3, "batched apply needs at least one vector");
  G2G_CHECK(b.X != nullptr, -4,
            "complex64 operator not made yet: make it with a complex64 apply or batch before the "
            "allocation-free apply");
  size_t need1 = 0;
  const int kcap = op_noalloc_batch_max<R>(op, b, &need1);
  if (kcap < 1) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "grid too large for the allocation-free apply: one vector needs %zu bytes of scratch, "
                  "the library holds %zu",
                  need1, (size_t)NOALLOC_SCRATCH_BYTES);
    throw Error(-3, msg);
  }
  G2G_CHECK(op.have_tensor && op.have_sigma, -4, "operator not ready: tensor or conductivity missing");
  G2G_CHECK(!op.prof.on, -3, "the profiling mode creates events; turn it off for the allocation-free apply");
  int dev = 0;
  G2G_HIP(hipGetDevice(&dev));
  G2G_CHECK(dev >= 0 && dev < MAX_DEVICES && scratch_dev[dev] != nullptr, -4,
            "no operator was created on the current device");
  // the operator's buffers with X and the spectrum slots moved into scratch;
  // the single-vector path's spectrum S is the first slot
  OpBuffers<R> bl = b;
  T2* base = static_cast<T2*>(scratch_dev[dev]);
  const size_t nX = (size_t)b.nchunk * b.NR * op.Ny * 2 * op.Nx;
  bl.X = base;
  bl.SB = base + nX;
  bl.S = bl.SB;
  bl.sb_cap = kcap;
  const int64_t nv = 3 * (int64_t)op.ncell();
  for (int r0 = 0; r0 < nrhs;) {
    const int k = std::min(kcap, nrhs - r0);
    if (k == 1) {
      op_apply<R>(op, bl, vin + nv * r0, vout + nv * r0, s, false, false);
    } else {
      const T2* vins[8];
      T2* vouts[8];
      for (int r = 0; r < k; r++) {
        vins[r] = vin + nv * (r0 + r);
        vouts[r] = vout + nv * (r0 + r);
      }
      op_apply_k<R>(op, bl, k, vins, vouts, s);
    }
    r0 += k;
  }
}
template int op_noalloc_batch_max<double>(const Operator&, const OpBuffers<double>&, size_t*);
template int op_noalloc_batch_max<float>(const Operator&, const OpBuffers<float>&, size_t*);
template void op_apply_batch_noalloc<double>(Operator&, OpBuffers<double>&, int, const zd*, zd*, hipStream_t);
template void op_apply_batch_noalloc<float>(Operator&, OpBuffers<float>&, int, const zf*, zf*, hipStream_t);

// ============================================================ receivers ==
// Continuation of the solution to receivers (ReCalculation, APPLY_RC_E_OP,
// APPLY_RC_H_OP, Continuation_Function_Module.F90:198-481): j = dsig * E is
// transformed once (the reference transforms it once for E and again for
// H); one contraction gives the six field components of every receiver; one
// inverse transform of 6 Nr rows brings them back.
//   K1 k_fwd_x         : j = dsig * E, zero pad, x-FFT (the IE kernel)
//   K2 k_fwd_y<FULL>   : y-FFT, the whole 2Nx x 2Ny spectrum
//   K3 k_rc_contract   : per (point, receiver) the 17 block products
//                        out_(c', r) = sum_z G_e(z, r) j_c(z) (six outputs)
//   K4 k_inv_y<FULL>   : inverse y-FFT of the 6 Nr output rows, y < Ny
//   K5 k_rc_inv_x      : inverse x-FFT, x < Nx, 1/(4 Nx Ny), (Nr,3,Nx,Ny) store

// One P-lane segment per (point, receiver) pair, lane = source layer z;
// partial products summed over the segment with lane exchanges.  All lanes
// stay active through the exchanges (loads and stores are predicated).
// The six output products of one (point, receiver) pair over the source
// layers z0 + lane (zero where that is past Nz or the pair is out of range).
__device__ inline void rc_products(const zd* __restrict__ S, const zd* __restrict__ G, int64_t pt, int r,
                                   int Nr, int Nz, int z, bool ok, zd& o0, zd& o1, zd& o2, zd& o3, zd& o4,
                                   zd& o5) {
  const zd* s = S + pt * 3 * Nz + z;
  const zd* g = G + (pt * 15 * Nr + r) * Nz + z;  // entry e at g[e * Nr * Nz]
  const int64_t es = (int64_t)Nr * Nz;
  // predicated loads into zero-initialised registers (a select between a
  // global element and a local constant would put the constant in scratch)
  zd jx = mkz(0, 0), jy = jx, jz = jx, rexx = jx, rexy = jx, rexz = jx, reyy = jx, reyz = jx,
     rezx = jx, rezy = jx, rezz = jx, rhxx = jx, rhxy = jx, rhxz = jx, rhyx = jx, rhyz = jx,
     rhzx = jx, rhzy = jx;
  if (ok) {
    jx = s[0];
    jy = s[Nz];
    jz = s[2 * Nz];
    rexx = ld_stream(g);
    rexy = ld_stream(g + es);
    rexz = ld_stream(g + 2 * es);
    reyy = ld_stream(g + 3 * es);
    reyz = ld_stream(g + 4 * es);
    rezx = ld_stream(g + 5 * es);
    rezy = ld_stream(g + 6 * es);
    rezz = ld_stream(g + 7 * es);
    rhxx = ld_stream(g + 8 * es);
    rhxy = ld_stream(g + 9 * es);
    rhxz = ld_stream(g + 10 * es);
    rhyx = ld_stream(g + 11 * es);
    rhyz = ld_stream(g + 12 * es);
    rhzx = ld_stream(g + 13 * es);
    rhzy = ld_stream(g + 14 * es);
  }
  // APPLY_RC_E_OP (:420-434) and APPLY_RC_H_OP (:461-473); RHYY = -RHXX (:468).
  // Named registers, not an array, so that nothing is placed in scratch.
  o0 = rexx * jx + rexy * jy + rexz * jz;
  o1 = rexy * jx + reyy * jy + reyz * jz;
  o2 = rezx * jx + rezy * jy + rezz * jz;
  o3 = rhxx * jx + rhxy * jy + rhxz * jz;
  o4 = rhyx * jx - rhxx * jy + rhyz * jz;
  o5 = rhzx * jx + rhzy * jy;
}

// CHUNKS: more than P (= 64) source layers; a lane then also takes the
// layers lane + P, lane + 2P, ... (one round otherwise, the case's own code)
template <int P, bool CHUNKS = false>
__global__ __launch_bounds__(NT) void k_rc_contract(const zd* __restrict__ S, const zd* __restrict__ G,
                                                    zd* __restrict__ So, int64_t npairs, int Nr, int Nz) {
  const int lane = threadIdx.x & (P - 1);
  const int64_t q = (int64_t)blockIdx.x * (NT / P) + threadIdx.x / P;
  const bool pair_ok = q < npairs;
  const int64_t pt = pair_ok ? q / Nr : 0;
  const int r = pair_ok ? (int)(q - pt * Nr) : 0;
  zd o0, o1, o2, o3, o4, o5;
  rc_products(S, G, pt, r, Nr, Nz, lane, pair_ok && lane < Nz, o0, o1, o2, o3, o4, o5);
  if constexpr (CHUNKS) {
    for (int z0 = P; z0 < Nz; z0 += P) {
      zd p0, p1, p2, p3, p4, p5;
      rc_products(S, G, pt, r, Nr, Nz, z0 + lane, pair_ok && z0 + lane < Nz, p0, p1, p2, p3, p4, p5);
      o0 = o0 + p0;
      o1 = o1 + p1;
      o2 = o2 + p2;
      o3 = o3 + p3;
      o4 = o4 + p4;
      o5 = o5 + p5;
    }
  }
#pragma unroll
  for (int off = P / 2; off > 0; off >>= 1) {
#define G2G_XR(v) v.x += __shfl_xor(v.x, off); v.y += __shfl_xor(v.y, off);
    G2G_XR(o0) G2G_XR(o1) G2G_XR(o2) G2G_XR(o3) G2G_XR(o4) G2G_XR(o5)
#undef G2G_XR
  }
  if (pair_ok && lane == 0) {
    zd* d = So + pt * 6 * Nr + r;
    d[0] = o0;
    d[Nr] = o1;
    d[2 * Nr] = o2;
    d[3 * Nr] = o3;
    d[4 * Nr] = o4;
    d[5 * Nr] = o5;
  }
}

// Inverse x-FFT of the output rows (row = component*Nr + receiver), keep
// x < Nx, scale (RC_OP_FFTW_BWD, :344-348) and store Ea, Ha (Nr,3,Nx,Ny).
template <int L, bool RT = false>
__global__ __launch_bounds__(NT) void k_rc_inv_x(const zd* __restrict__ X, zd* __restrict__ Ea,
                                                 zd* __restrict__ Ha, const zd* __restrict__ twg, Geo g,
                                                 int Nr, double scale) {
  constexpr int NR = nr_for<double>(L);
  constexpr int LS = L + 1;
  __shared__ zd buf[NR * LS];
  __shared__ zd tw[L];
  const int y = blockIdx.x, chunk = blockIdx.y, Nx = g.Nx;
  const int Lr = RT ? 2 * Nx : L;
  load_tw<zd>(tw, twg, Lr);
  const zd* src = X + (((int64_t)chunk * g.Ny + y) * Lr) * NR;
  for (int i = threadIdx.x; i < NR * Lr; i += NT) {
    const int kx = i / NR, s = i - kx * NR;
    buf[s * LS + kx] = src[i];
  }
  __syncthreads();
  lds_fft_any<zd, L, NR, LS, NT, +1, RT>(buf, tw, g.px);
  for (int i = threadIdx.x; i < NR * Nx; i += NT) {
    const int s = i / Nx, x = i - s * Nx, row = chunk * NR + s;
    if (row < g.Nz3) {
      const int c = row / Nr, r = row - c * Nr;
      const int64_t o = r + (int64_t)Nr * ((c % 3) + 3 * (x + (int64_t)Nx * y));
      (c < 3 ? Ea : Ha)[o] = buf[s * LS + x] * scale;
    }
  }
}

// LoadEint (:255): dsig = sigma_a - sigma_b
__global__ void k_dsig(const zd* __restrict__ csiga, const zd* __restrict__ csigb, zd* __restrict__ dsig,
                       int64_t ncxy, int Nz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncxy * Nz) return;
  const int z = (int)(i / ncxy);
  dsig[i] = csiga[i] - csigb[z];
}

void rc_init(RcOperator& rc, Operator& op, int Nr, const double* xs, const double* ys, const double* depth) {
  G2G_CHECK(Nr >= 1 && Nr <= 64, -3, "Nr must be in [1,64]");
  for (int r = 0; r < Nr; r++) {
    G2G_CHECK(std::isfinite(xs[r]) && std::isfinite(ys[r]) && std::isfinite(depth[r]) && depth[r] >= 0, -3,
              "receiver shifts must be finite and depths non-negative");
    // the weights need a nonzero distance to the source cell corner
    // (VBTransformWeightsAllSingleInt takes log(rho))
    const double fx = xs[r] / op.dx - 0.5, fy = ys[r] / op.dy - 0.5;
    const bool cx = std::fabs(fx - std::nearbyint(fx)) < 1e-12 && std::fabs(std::nearbyint(fx)) < op.Nx;
    const bool cy = std::fabs(fy - std::nearbyint(fy)) < 1e-12 && std::fabs(std::nearbyint(fy)) < op.Ny;
    G2G_CHECK(!(cx && cy), -3, "receiver on a cell corner: the filter weights are singular there");
  }
  rc.op = &op;
  rc.Nr = Nr;
  rc.xs.assign(xs, xs + Nr);
  rc.ys.assign(ys, ys + Nr);
  rc.depth.assign(depth, depth + Nr);
  const int Nx2 = 2 * op.Nx, Ny2 = 2 * op.Ny;
  rc.NP = (int64_t)Nx2 * Ny2;
  rc.BS = 15 * (int64_t)Nr * op.Nz;
  rc.NR = nr_rt<double>(Nx2);
  rc.nchunk_f = (3 * op.Nz + rc.NR - 1) / rc.NR;
  rc.nchunk_i = (6 * Nr + rc.NR - 1) / rc.NR;
  auto A = [&](zd** p, size_t n) {
    G2G_HIP(hipMalloc((void**)p, n * sizeof(zd)));
    rc.bytes += n * sizeof(zd);
  };
  A(&rc.G, (size_t)rc.NP * rc.BS);
  A(&rc.dsig, (size_t)op.ncell());
  A(&rc.Xf, (size_t)rc.nchunk_f * rc.NR * op.Ny * Nx2);
  A(&rc.S, (size_t)rc.NP * 3 * op.Nz);
  A(&rc.So, (size_t)rc.NP * 6 * Nr);
  A(&rc.Xi, (size_t)rc.nchunk_i * rc.NR * op.Ny * Nx2);
}

void rc_fini(RcOperator& rc) {
  for (zd** p : {&rc.G, &rc.dsig, &rc.Xf, &rc.S, &rc.So, &rc.Xi, &rc.tabs}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  if (rc.wts) (void)hipFree(rc.wts);
  rc.wts = nullptr;
}

// FinalizeRCOperator (:168-196): the conductivity of the current frequency
void rc_set_sigma(RcOperator& rc, const zd* d_csiga) {
  Operator& op = *rc.op;
  G2G_CHECK(op.have_bkg, -4, "background not set");
  zd* d_csigb;
  G2G_HIP(hipMalloc(&d_csigb, sizeof(zd) * op.Nz));
  G2G_HIP(hipMemcpy(d_csigb, op.csigb.data(), sizeof(zd) * op.Nz, hipMemcpyHostToDevice));
  const int64_t n = op.ncell();
  hipLaunchKernelGGL(k_dsig, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, op.stream, d_csiga, d_csigb,
                     rc.dsig, (int64_t)op.Nx * op.Ny, op.Nz);
  G2G_HIP(hipGetLastError());
  G2G_HIP(hipStreamSynchronize(op.stream));
  (void)hipFree(d_csigb);
  rc.have_sigma = true;
  rc.sigma_version = op.bkg_version;
}

void rc_apply(RcOperator& rc, const zd* E, zd* Ea, zd* Ha, hipStream_t s) {
  Operator& op = *rc.op;
  G2G_CHECK(rc.have_tensor && rc.spectral && rc.tensor_version == op.bkg_version, -4,
            "receiver tensor not built and transformed for the current frequency");
  G2G_CHECK(rc.have_sigma && rc.sigma_version == op.bkg_version, -4,
            "receiver conductivity not set for the current frequency");
  const int Nx2 = 2 * op.Nx, Ny2 = 2 * op.Ny, Nr = rc.Nr;
  Geo gf = make_geo(op, rc.NR);
  Geo gi = gf;
  gi.Nz3 = 6 * Nr;
  const int nry = nr_rt<double>(Ny2);
  const zd* twx = op.d64.twx;
  const zd* twy = op.d64.twy;
  G2G_HIP(hipEventRecord(op.ev[4], s));
  len_dispatch(Nx2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    hipLaunchKernelGGL((k_fwd_x<double, LL, decltype(LC)::rt>), dim3(op.Ny, rc.nchunk_f), dim3(NT), 0, s, E,
                       rc.dsig, rc.Xf, twx, gf);
  });
  len_dispatch(Ny2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    hipLaunchKernelGGL((k_fwd_y<double, LL, true, decltype(LC)::rt>), dim3(Nx2, (3 * op.Nz + nry - 1) / nry),
                       dim3(NT), 0, s, rc.Xf, rc.S, twy, gf);
  });
  G2G_HIP(hipEventRecord(op.ev[7], s));
  {
    int P = 1;
    while (P < op.Nz && P < 64) P <<= 1;  // above 64 layers a lane takes several
    const int64_t npairs = rc.NP * Nr;
    const unsigned blocks = (unsigned)((npairs + NT / P - 1) / (NT / P));
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(blocks), dim3(NT), 0, s, rc.S, rc.G, rc.So, npairs, Nr, op.Nz);
    };
    switch (P) {
      case 1: go(k_rc_contract<1>); break;
      case 2: go(k_rc_contract<2>); break;
      case 4: go(k_rc_contract<4>); break;
      case 8: go(k_rc_contract<8>); break;
      case 16: go(k_rc_contract<16>); break;
      case 32: go(k_rc_contract<32>); break;
      default:
        if (op.Nz > 64) go(k_rc_contract<64, true>);
        else go(k_rc_contract<64>);
        break;
    }
  }
  G2G_HIP(hipEventRecord(op.ev[5], s));
  len_dispatch(Ny2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    hipLaunchKernelGGL((k_inv_y<double, LL, true, decltype(LC)::rt>), dim3(Nx2, (6 * Nr + nry - 1) / nry),
                       dim3(NT), 0, s, rc.So, rc.Xi, twy, gi);
  });
  len_dispatch(Nx2, [&](auto LC) {
    constexpr int LL = decltype(LC)::value;
    hipLaunchKernelGGL((k_rc_inv_x<LL, decltype(LC)::rt>), dim3(op.Ny, rc.nchunk_i), dim3(NT), 0, s, rc.Xi, Ea,
                       Ha, twx, gi, Nr, 1.0 / (4.0 * op.Nx * op.Ny));
  });
  G2G_HIP(hipEventRecord(op.ev[6], s));
  G2G_HIP(hipGetLastError());
  rc.applies++;
}

// ------------------------------------------------------ conductivity ---
template <typename R>
__global__ void k_sigma(const zd* __restrict__ csiga, const zd* __restrict__ csigb,
                        const double* __restrict__ sq, cx_t<R>* gsig, cx_t<R>* asiga,
                        int64_t ncxy, int Nz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncxy * Nz) return;
  const int z = (int)(i / ncxy);
  const zd sa = csiga[i], sb = csigb[z];
  // a = 2 sqrt(Re sb) / (sa + conj(sb)); g = (sa - sb) * a
  const zd a = cdiv(mkz(2.0 * sq[z], 0.0), mkz(sa.x + sb.x, sa.y - sb.y));
  const zd gg = mkz(sa.x - sb.x, sa.y - sb.y) * a;
  gsig[i] = mk<cx_t<R>>((R)gg.x, (R)gg.y);
  asiga[i] = mk<cx_t<R>>((R)a.x, (R)a.y);
}

template <typename R> __global__ void k_to_real(const double* a, R* b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (R)a[i];
}

__global__ void k_z_to_f(const zd* __restrict__ a, zf* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = zf{(float)a[i].x, (float)a[i].y};
}

// complex64 copies of the per-cell scalings for the complex64 operator
void op_sigma_to_f32(Operator& op) {
  const int64_t n = op.ncell();
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_z_to_f, dim3(blocks), dim3(256), 0, op.stream, op.d64.gsig, op.d32.gsig, n);
  hipLaunchKernelGGL(k_z_to_f, dim3(blocks), dim3(256), 0, op.stream, op.d64.asiga, op.d32.asiga, n);
  hipLaunchKernelGGL((k_to_real<float>), dim3(1), dim3(256), 0, op.stream, op.d64.sq, op.d32.sq, op.Nz);
  G2G_HIP(hipGetLastError());
}

// ------------------------------------------------------------ adjoint ---
// The contraction kernel K of A = diag(r) (diag(a) - K diag(g)) satisfies
// diag(dz) K = (diag(dz) K)^T (reciprocity of the Galerkin matrix,
// giem2g.pdf Eq. 20), so A^T = diag(a r) - diag(g dz) K diag(r/dz) and
// A^H v = conj(A^T conj(v)).  K u is one operator pass with the per-cell
// pre-scale r/dz, no diagonal term and unit post-scale.
__global__ void k_adj_pre(const zd* __restrict__ v, const double* __restrict__ sq,
                          const double* __restrict__ dz, int64_t ncxy, int Nz, zd* __restrict__ u,
                          zd* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = ncxy * Nz;
  if (i >= 3 * nc) return;
  const int z = (int)((i % nc) / ncxy);
  u[i] = mkz(v[i].x, -v[i].y);  // conj(v)
  if (i < nc) w[i] = mkz(sq[z] / dz[z], 0.0);
}

__global__ void k_adj_post(const zd* __restrict__ u, const zd* __restrict__ kv,
                           const zd* __restrict__ asiga, const zd* __restrict__ gsig,
                           const double* __restrict__ sq, const double* __restrict__ dz, int64_t ncxy,
                           int Nz, zd* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = ncxy * Nz;
  if (i >= 3 * nc) return;
  const int64_t cell = i % nc;
  const int z = (int)(cell / ncxy);
  // kv = -K(r u / dz): A^T u = a r u + g dz kv
  const zd t = asiga[cell] * u[i] * sq[z] + gsig[cell] * kv[i] * dz[z];
  out[i] = mkz(t.x, -t.y);
}

__global__ void k_fill_z(zd* __restrict__ a, zd val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = val;
}

void op_apply_adjoint(Operator& op, const zd* vin, zd* vout, hipStream_t s) {
  G2G_CHECK(op.have_tensor && op.have_sigma, -4, "operator not ready: tensor or conductivity missing");
  const int64_t nc = op.ncell(), nv = op.nvec();
  if (!op.adj) {
    G2G_HIP(hipMalloc(&op.adj, sizeof(zd) * (2 * nv + 2 * nc) + sizeof(double) * 2 * op.Nz));
  }
  zd* u = op.adj;
  zd* kv = u + nv;
  zd* w = kv + nv;
  zd* zero_c = w + nc;
  double* one_z = reinterpret_cast<double*>(zero_c + nc);
  double* dzd = one_z + op.Nz;
  {
    std::vector<double> h(2 * op.Nz);
    for (int z = 0; z < op.Nz; z++) {
      h[z] = 1.0;
      h[op.Nz + z] = op.dz[z];
    }
    G2G_HIP(hipMemcpyAsync(one_z, h.data(), sizeof(double) * 2 * op.Nz, hipMemcpyHostToDevice, s));
  }
  const unsigned nb = (unsigned)((3 * nc + 255) / 256);
  hipLaunchKernelGGL(k_fill_z, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, zero_c, mkz(0.0, 0.0), nc);
  hipLaunchKernelGGL(k_adj_pre, dim3(nb), dim3(256), 0, s, vin, op.d64.sq, dzd, (int64_t)op.Nx * op.Ny, op.Nz,
                     u, w);
  OpBuffers<double> b = op.d64;  // shares the tensor, spectra and twiddles
  b.gsig = w;
  b.asiga = zero_c;
  b.sq = one_z;
  op_apply<double>(op, b, u, kv, s, false);
  hipLaunchKernelGGL(k_adj_post, dim3(nb), dim3(256), 0, s, u, kv, op.d64.asiga, op.d64.gsig, op.d64.sq, dzd,
                     (int64_t)op.Nx * op.Ny, op.Nz, vout);
  G2G_HIP(hipGetLastError());
  G2G_HIP(hipStreamSynchronize(s));
}

void op_set_sigma(Operator& op, const zd* d_csiga) {
  G2G_CHECK(op.have_bkg, -4, "background not set");
  const int Nz = op.Nz;
  if (!op.csiga) G2G_HIP(hipMalloc(&op.csiga, sizeof(zd) * op.ncell()));
  if (d_csiga != op.csiga)
    G2G_HIP(hipMemcpyAsync(op.csiga, d_csiga, sizeof(zd) * op.ncell(), hipMemcpyDeviceToDevice, op.stream));
  zd* d_csigb;
  double* d_sq;
  G2G_HIP(hipMalloc(&d_csigb, sizeof(zd) * Nz));
  G2G_HIP(hipMalloc(&d_sq, sizeof(double) * Nz));
  G2G_HIP(hipMemcpy(d_csigb, op.csigb.data(), sizeof(zd) * Nz, hipMemcpyHostToDevice));
  G2G_HIP(hipMemcpy(d_sq, op.sqsigb.data(), sizeof(double) * Nz, hipMemcpyHostToDevice));
  const int64_t n = op.ncell(), ncxy = (int64_t)op.Nx * op.Ny;
  const int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL((k_sigma<double>), dim3(blocks), dim3(256), 0, op.stream, op.csiga, d_csigb,
                     d_sq, op.d64.gsig, op.d64.asiga, ncxy, Nz);
  G2G_HIP(hipMemcpyAsync(op.d64.sq, d_sq, sizeof(double) * Nz, hipMemcpyDeviceToDevice, op.stream));
  op.have_sigma = true;
  if (op.d32.gsig) op_sigma_to_f32(op);
  G2G_HIP(hipGetLastError());
  G2G_HIP(hipStreamSynchronize(op.stream));
  (void)hipFree(d_csigb);
  (void)hipFree(d_sq);
}

template <typename R> static void upload_dzn(const Operator& op, OpBuffers<R>& b) {
  if (!b.dzn) return;
  std::vector<R> dzn(op.Nz);
  for (int z = 0; z < op.Nz; z++) dzn[z] = (R)(1.0 / (op.dz[z] * 4.0 * op.Nx * op.Ny));
  G2G_HIP(hipMemcpy(b.dzn, dzn.data(), sizeof(R) * op.Nz, hipMemcpyHostToDevice));
}

void op_set_geometry(Operator& op, double dx, double dy, const double* z) {
  G2G_CHECK(dx > 0 && dy > 0, -3, "dx, dy must be positive");
  std::vector<double> dz(op.Nz);
  for (int i = 0; i < op.Nz; i++) {
    dz[i] = z[i + 1] - z[i];
    G2G_CHECK(dz[i] > 0, -3, "z borders must increase");
  }
  G2G_HIP(hipStreamSynchronize(op.stream));
  if (dx != op.dx || dy != op.dy) op.have_wts = false;
  op.tensor_own = false;  // a tensor from the earlier geometry: no mirror until the next build
  op.dx = dx;
  op.dy = dy;
  op.z.assign(z, z + op.Nz + 1);
  op.dz = dz;
  upload_dzn<double>(op, op.d64);
  upload_dzn<float>(op, op.d32);
}

__global__ void k_g_to_f32(const zd* __restrict__ a, zf* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = zf{(float)a[i].x, (float)a[i].y};
}

void op_make_g32(Operator& op) {
  G2G_CHECK(op.have_tensor, -4, "tensor not computed");
  if (!op.d32.X) op_alloc_buffers<float>(op, op.d32, true);
  const int64_t n = (int64_t)op.Nx * op.Ny * op.BS;
  hipLaunchKernelGGL(k_g_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, op.stream,
                     op.d64.G, op.d32.G, n);
  G2G_HIP(hipGetLastError());
  if (op.have_sigma) op_sigma_to_f32(op);
  G2G_HIP(hipStreamSynchronize(op.stream));
  op.have_g32 = true;
}

template void op_alloc_buffers<double>(Operator&, OpBuffers<double>&, bool);
template void op_alloc_buffers<float>(Operator&, OpBuffers<float>&, bool);
template void op_free_buffers<double>(OpBuffers<double>&);
template void op_free_buffers<float>(OpBuffers<float>&);
template void op_apply<double>(Operator&, OpBuffers<double>&, const zd*, zd*,
                               hipStream_t, bool, bool);
template void op_apply<float>(Operator&, OpBuffers<float>&, const zf*, zf*, hipStream_t,
                              bool, bool);

}  // namespace g2g
