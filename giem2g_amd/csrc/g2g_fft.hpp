// g2g_fft.hpp -- batched power-of-two complex FFTs held in LDS.
//
// NR independent rows of length L sit in LDS at buf[row*LS + i] (LS >= L, a
// padded row stride).  The transform is an in-place Stockham autosort FFT
// with radix-4 stages (plus one radix-2 stage when log2 L is odd): every
// thread loads its butterflies into registers, the block synchronises, and
// the outputs are written back to the same buffer.  Twiddles come from an
// LDS table tw[t] = exp(-2*pi*i*t/L).  DIR = -1 is the forward transform
// (exp(-i)), DIR = +1 the unnormalised inverse.
#pragma once
#include "g2g_common.hpp"

namespace g2g {

template <int R, int DIR, typename T2> __device__ inline void dft_small(T2* v) {
  if constexpr (R == 2) {
    T2 a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else {
    T2 a0 = v[0] + v[2], a1 = v[0] - v[2];
    T2 a2 = v[1] + v[3], d = v[1] - v[3];
    T2 a3 = (DIR < 0) ? mul_mi(d) : mul_pi(d);
    v[0] = a0 + a2;
    v[2] = a0 - a2;
    v[1] = a1 + a3;
    v[3] = a1 - a3;
  }
}

template <typename T2, int L, int NR, int LS, int NT, int R, int DIR>
__device__ inline void stockham_stage(T2* buf, const T2* tw, int Ns) {
  constexpr int NBR = L / R;
  constexpr int NB = NR * NBR;
  constexpr int PER = (NB + NT - 1) / NT;
  T2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int bi = tid + q * NT;
    if (NB % NT == 0 || bi < NB) {
      const int row = bi / NBR, j = bi % NBR;
      const int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; r++) v[q][r] = buf[row * LS + j + r * NBR];
      if (Ns > 1) {
        const int tstep = L / (Ns * R);
#pragma unroll
        for (int r = 1; r < R; r++) {
          T2 w = tw[k * r * tstep];
          if (DIR > 0) w.y = -w.y;
          v[q][r] = v[q][r] * w;
        }
      }
      dft_small<R, DIR>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int bi = tid + q * NT;
    if (NB % NT == 0 || bi < NB) {
      const int row = bi / NBR, j = bi % NBR;
      const int k = j & (Ns - 1);
      const int idxD = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) buf[row * LS + idxD + r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

// Full transform.  Caller must __syncthreads() after filling buf and tw.
template <typename T2, int L, int NR, int LS, int NT, int DIR>
__device__ inline void lds_fft(T2* buf, const T2* tw) {
  constexpr int LOG = (L <= 1) ? 0 : (31 - __builtin_clz(L));
  constexpr int N4 = LOG / 2;
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < N4; s++) {
    stockham_stage<T2, L, NR, LS, NT, 4, DIR>(buf, tw, Ns);
    Ns *= 4;
  }
  if constexpr (LOG & 1) stockham_stage<T2, L, NR, LS, NT, 2, DIR>(buf, tw, Ns);
}

}  // namespace g2g
