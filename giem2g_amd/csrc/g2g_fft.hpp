// g2g_fft.hpp -- batched complex FFTs held in LDS: radix-4 Stockham for
// powers of two, and a factor-plan Stockham for other lengths (below).
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

// ------------------------------------------------------- any length ------
// Lengths that are not powers of two (2Nx, 2Ny for any even Nx, Ny) run a
// Stockham autosort FFT over the length's factors -- fours, then twos, then
// the odd primes -- in the same LDS layout.  A stage of radix R computes
// every output as the R-term sum over its butterfly's inputs, each times one
// entry of the twiddle table tw[t] = exp(-2 pi i t/L): output slot
// t = b Ns R + r Ns + k of a row takes the inputs j + q L/R, j = b Ns + k,
// with factors tw[(q e) mod L], e = k L/(Ns R) + r L/R (the stage twiddle
// w_{Ns R}^{q k} and the radix-R DFT in one).  That is O(L R) work per
// stage and row for any prime R.  The plan is made on the host.
struct FftPlan {
  int L = 0, nst = 0;
  int rad[24] = {};
};

inline FftPlan fft_plan(int L) {
  FftPlan p;
  p.L = L;
  int n = L;
  auto push = [&](int r) {
    if (p.nst >= 24) throw Error(-3, "FFT length has too many factors");
    p.rad[p.nst++] = r;
    n /= r;
  };
  while (n % 4 == 0) push(4);
  while (n % 2 == 0) push(2);
  for (int f = 3; n > 1; f += 2)
    while (n % f == 0) push(f);
  return p;
}

template <typename T2, int PER, int NT, int DIR>
__device__ inline void stockham_stage_any(T2* buf, const T2* tw, int L, int NR, int LS, int R, int Ns) {
  const int NBR = L / R, NO = NR * L, sK = L / (Ns * R);
  T2 o[PER];
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int oi = threadIdx.x + p * NT;
    o[p] = mk<T2>(0, 0);
    if (oi < NO) {
      const int row = oi / L, t = oi - row * L;
      const int k = t % Ns, r = (t / Ns) % R, b = t / (Ns * R);
      const T2* src = buf + row * LS + b * Ns + k;
      const int e = k * sK + r * NBR;
      T2 acc = src[0];
      int idx = e;
      for (int q = 1; q < R; q++) {
        T2 w = tw[idx];
        if (DIR > 0) w.y = -w.y;
        acc = acc + src[q * NBR] * w;
        idx += e;
        if (idx >= L) idx -= L;
      }
      o[p] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int oi = threadIdx.x + p * NT;
    if (oi < NO) {
      const int row = oi / L;
      buf[row * LS + (oi - row * L)] = o[p];
    }
  }
  __syncthreads();
}

// Full transform of NR rows of length plan.L <= LC (the capacity the
// buffers were sized for).  RT = false: LC itself, a power of two, with
// lds_fft.  Caller must __syncthreads() after filling buf and tw.
template <typename T2, int LC, int NR, int LS, int NT, int DIR, bool RT>
__device__ inline void lds_fft_any(T2* buf, const T2* tw, const FftPlan& plan) {
  if constexpr (!RT) {
    lds_fft<T2, LC, NR, LS, NT, DIR>(buf, tw);
  } else {
    constexpr int PER = (NR * LC + NT - 1) / NT;
    int Ns = 1;
    for (int s = 0; s < plan.nst; s++) {
      stockham_stage_any<T2, PER, NT, DIR>(buf, tw, plan.L, NR, LS, plan.rad[s], Ns);
      Ns *= plan.rad[s];
    }
  }
}

}  // namespace g2g
