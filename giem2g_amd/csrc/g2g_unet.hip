// g2g_unet.hip -- the learned right preconditioner on MI355X (gfx950).
//
// Network: 3-level 3D U-Net, widths 16/32/64, input 8 channels (residual
// re/im x 3, log10 contrast, log10 of the cell size in background
// wavenumbers), output 6 channels; 3x3x3
// convolutions with bias and ReLU, 2x2x2 average pooling, nearest
// upsampling, concatenation [upsampled, skip], a 1x1 output convolution.
// Volumes are channels-last [batch][z][y][x][C] in bf16.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   out[co][p] = bias[co] + sum_k W[co][k] X[k][p],  k = tap*Cin + ci,
// A = weights (16 output channels x 32 k), B = input (32 k x 16 voxels).
// Lane l of a wave holds A row l%16 / B column l%16 and the 8 consecutive
// k = 8*(l/16) .. 8*(l/16)+7, which is one 16-byte load of 8 channels of one
// tap (Cin is a multiple of 8); the FP32 result of lane l is rows
// 4*(l/16) .. 4*(l/16)+3 of column l%16 -- four consecutive output channels
// of one voxel, stored as one 8-byte bf16 write.  The data gradient is the
// same kernel run on the transposed, tap-flipped weights; the weight
// gradient accumulates in FP32 through LDS tiles.  No reference counterpart:
// GFGMRES/seq_fgmres.F90:142-152 preconditions with an inner GMRES.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "g2g_op.hpp"
#include "g2g_unet.hpp"

namespace g2g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int UB = 256;

int64_t up_to(int64_t a, int64_t m) { return (a + m - 1) / m * m; }
unsigned nblk(int64_t n, int b = UB) { return (unsigned)((n + b - 1) / b); }

struct Src {
  const bf16* p;
  int C;
  int up;  // 1: the source is at half resolution (nearest upsampling)
  int64_t bstride;
};

struct ConvArgs {
  Src s1, s2;
  const bf16* w;
  int kp, ntap, cin;
  const float* bias;
  int cout, relu;
  bf16* ybf;
  float* yf;
  int ldy;
  int64_t ybstride;
  int nx, ny, nz;
};

// One wave: 16 output channels x 64 voxels (4 MFMA columns); 4 waves per block.
__global__ __launch_bounds__(UB) void k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int nvox = a.nx * a.ny * a.nz;
  const int pbase = (blockIdx.x * 4 + wave) * 64;
  const int cb = blockIdx.y, bat = blockIdx.z;
  const bf16* x1 = a.s1.p + bat * a.s1.bstride;
  const bf16* x2 = a.s2.p ? a.s2.p + bat * a.s2.bstride : a.s1.p;
  int vx0, vy0, vz0, vx1, vy1, vz1, vx2, vy2, vz2, vx3, vy3, vz3;
#define G2G_VXYZ(s)                                        \
  {                                                        \
    int p_ = pbase + (s) * 16 + col;                       \
    if (p_ >= nvox) p_ = -1;                               \
    int q_ = p_ < 0 ? 0 : p_;                              \
    vx##s = q_ % a.nx;                                     \
    q_ /= a.nx;                                            \
    vy##s = q_ % a.ny;                                     \
    vz##s = p_ < 0 ? -1000 : q_ / a.ny;                    \
  }
  G2G_VXYZ(0) G2G_VXYZ(1) G2G_VXYZ(2) G2G_VXYZ(3)
#undef G2G_VXYZ
  f32x4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  const bf16* wrow = a.w + (int64_t)(cb * 16 + col) * a.kp;
  const int nks = a.kp / 32;
  const int nxh = a.nx >> 1, nyh = a.ny >> 1;
  for (int ks = 0; ks < nks; ks++) {
    const int k0 = ks * 32 + kq * 8;
    const bf16x8 af = *reinterpret_cast<const bf16x8*>(wrow + k0);
    const int tap = k0 / a.cin;
    const int ci = k0 - tap * a.cin;
    const bool tv = tap < a.ntap;
    int dx = 0, dy = 0, dz = 0;
    if (a.ntap == 27) {
      dz = tap / 9 - 1;
      dy = (tap / 3) % 3 - 1;
      dx = tap % 3 - 1;
    }
    const bool first = ci < a.s1.C;
    const bf16* sp = x1;
    int C = a.s1.C, cc = ci, up = a.s1.up;
    if (!first) {
      sp = x2;
      C = a.s2.C;
      cc = ci - a.s1.C;
      up = a.s2.up;
    }
#define G2G_BFRAG(s)                                                                    \
  {                                                                                     \
    bf16x8 bfr = {0, 0, 0, 0, 0, 0, 0, 0};                                              \
    const int X = vx##s + dx, Y = vy##s + dy, Z = vz##s + dz;                           \
    if (tv && X >= 0 && X < a.nx && Y >= 0 && Y < a.ny && Z >= 0 && Z < a.nz) {         \
      int64_t q;                                                                        \
      if (up) q = ((int64_t)(Z >> 1) * nyh + (Y >> 1)) * nxh + (X >> 1);               \
      else q = ((int64_t)Z * a.ny + Y) * a.nx + X;                                      \
      bfr = *reinterpret_cast<const bf16x8*>(sp + q * C + cc);                          \
    }                                                                                   \
    c##s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr, c##s, 0, 0, 0);             \
  }
    G2G_BFRAG(0) G2G_BFRAG(1) G2G_BFRAG(2) G2G_BFRAG(3)
#undef G2G_BFRAG
  }
  const int co0 = cb * 16 + kq * 4;
  float b0 = 0, b1 = 0, b2 = 0, b3 = 0;
  if (a.bias) {
    if (co0 < a.cout) b0 = a.bias[co0];
    if (co0 + 1 < a.cout) b1 = a.bias[co0 + 1];
    if (co0 + 2 < a.cout) b2 = a.bias[co0 + 2];
    if (co0 + 3 < a.cout) b3 = a.bias[co0 + 3];
  }
#define G2G_EPI(s)                                                                   \
  {                                                                                  \
    const int p = pbase + (s) * 16 + col;                                            \
    if (p < nvox && co0 < a.cout) {                                                  \
      float o0 = c##s[0] + b0, o1 = c##s[1] + b1, o2 = c##s[2] + b2, o3 = c##s[3] + b3; \
      if (a.relu) {                                                                  \
        o0 = fmaxf(o0, 0.f);                                                         \
        o1 = fmaxf(o1, 0.f);                                                         \
        o2 = fmaxf(o2, 0.f);                                                         \
        o3 = fmaxf(o3, 0.f);                                                         \
      }                                                                              \
      if (a.ybf) {                                                                   \
        bf16x4 o = {(bf16)o0, (bf16)o1, (bf16)o2, (bf16)o3};                         \
        *reinterpret_cast<bf16x4*>(a.ybf + bat * a.ybstride + (int64_t)p * a.ldy + co0) = o; \
      } else {                                                                       \
        float* y = a.yf + bat * a.ybstride + (int64_t)p * a.ldy + co0;               \
        y[0] = o0;                                                                   \
        if (co0 + 1 < a.cout) y[1] = o1;                                             \
        if (co0 + 2 < a.cout) y[2] = o2;                                             \
        if (co0 + 3 < a.cout) y[3] = o3;                                             \
      }                                                                              \
    }                                                                                \
  }
  G2G_EPI(0) G2G_EPI(1) G2G_EPI(2) G2G_EPI(3)
#undef G2G_EPI
}

// Weight gradient dW[co][t][ci] += sum_p dZ[p][co] X[p+off(t)][ci] and
// db[co] += sum_p dZ[p][co], FP32.  Block: one tap, a run of 64-voxel
// chunks; thread: a 4x4 (co, ci) register tile.
struct DwArgs {
  const float* dz;
  int ldz, cout, cout4;
  Src s1, s2;
  int cin, ntap;
  float *dW, *db;
  int nx, ny, nz;
  int64_t dzbstride;
  int chunks;
};

constexpr int DWV = 64;

__global__ __launch_bounds__(UB) void k_dw(DwArgs a) {
  extern __shared__ float sm[];
  float* sdz = sm;                    // [DWV][cout4]
  float* sx = sm + DWV * a.cout4;     // [DWV][cin]
  const int t = blockIdx.y, bat = blockIdx.z;
  const int nvox = a.nx * a.ny * a.nz;
  int dx = 0, dy = 0, dz = 0;
  if (a.ntap == 27) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
  }
  const int nci4 = a.cin / 4;
  const int tiles = (a.cout4 / 4) * nci4;
  const int tile = threadIdx.x;
  const int co = (tile / nci4) * 4, ci = (tile % nci4) * 4;
  float acc[16];
#pragma unroll
  for (int q = 0; q < 16; q++) acc[q] = 0.f;
  float bacc = 0.f;
  const float* dzb = a.dz + bat * a.dzbstride;
  const bf16* x1 = a.s1.p + bat * a.s1.bstride;
  const bf16* x2 = a.s2.p ? a.s2.p + bat * a.s2.bstride : a.s1.p;
  const int nxh = a.nx >> 1, nyh = a.ny >> 1;
  for (int ch = 0; ch < a.chunks; ch++) {
    const int p0 = (blockIdx.x * a.chunks + ch) * DWV;
    if (p0 >= nvox) break;
    for (int e = threadIdx.x; e < DWV * a.cout4; e += UB) {
      const int v = e / a.cout4, c = e - v * a.cout4;
      const int p = p0 + v;
      sdz[e] = (p < nvox && c < a.cout) ? dzb[(int64_t)p * a.ldz + c] : 0.f;
    }
    for (int e = threadIdx.x; e < DWV * a.cin; e += UB) {
      const int v = e / a.cin, c = e - v * a.cin;
      const int p = p0 + v;
      float val = 0.f;
      if (p < nvox) {
        const int x = p % a.nx, y = (p / a.nx) % a.ny, z = p / (a.nx * a.ny);
        const int X = x + dx, Y = y + dy, Z = z + dz;
        if (X >= 0 && X < a.nx && Y >= 0 && Y < a.ny && Z >= 0 && Z < a.nz) {
          const bf16* sp = x1;
          int C = a.s1.C, cc = c, up = a.s1.up;
          if (c >= a.s1.C) {
            sp = x2;
            C = a.s2.C;
            cc = c - a.s1.C;
            up = a.s2.up;
          }
          int64_t q;
          if (up) q = ((int64_t)(Z >> 1) * nyh + (Y >> 1)) * nxh + (X >> 1);
          else q = ((int64_t)Z * a.ny + Y) * a.nx + X;
          val = (float)sp[q * C + cc];
        }
      }
      sx[e] = val;
    }
    __syncthreads();
    if (tile < tiles) {
      for (int v = 0; v < DWV; v++) {
        const f32x4 g = *reinterpret_cast<const f32x4*>(sdz + v * a.cout4 + co);
        const f32x4 xv = *reinterpret_cast<const f32x4*>(sx + v * a.cin + ci);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[i * 4 + j] += g[i] * xv[j];
      }
    }
    if (t == 0 && threadIdx.x < a.cout)
      for (int v = 0; v < DWV; v++) bacc += sdz[v * a.cout4 + threadIdx.x];
    __syncthreads();
  }
  if (tile < tiles) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (co + i < a.cout) atomicAdd(a.dW + ((int64_t)(co + i) * a.ntap + t) * a.cin + ci + j, acc[i * 4 + j]);
  }
  if (t == 0 && threadIdx.x < a.cout) atomicAdd(a.db + threadIdx.x, bacc);
}

// forward weights w[co][tap*cin+ci] and transposed weights
// wt[ci][tap'*coutT+co] = W[co][ntap-1-tap'][ci] (bf16, zero padded)
__global__ void k_wprep(const float* __restrict__ W, int cout, int cin, int ntap, bf16* w, int coutP,
                        int kp, bf16* wt, int cinP, int coutT, int kpT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n1 = (int64_t)coutP * kp, n2 = (int64_t)cinP * kpT;
  if (i < n1) {
    const int co = (int)(i / kp), k = (int)(i % kp);
    const int t = k / cin, ci = k % cin;
    float v = 0.f;
    if (co < cout && t < ntap) v = W[((int64_t)co * ntap + t) * cin + ci];
    w[i] = (bf16)v;
  } else if (i < n1 + n2) {
    const int64_t j = i - n1;
    const int ci = (int)(j / kpT), k = (int)(j % kpT);
    const int tp = k / coutT, co = k % coutT;
    float v = 0.f;
    if (ci < cin && tp < ntap && co < cout) v = W[((int64_t)co * ntap + (ntap - 1 - tp)) * cin + ci];
    wt[j] = (bf16)v;
  }
}

// 2x2x2 average pooling, 8 channels per thread
__global__ void k_pool(const bf16* __restrict__ src, bf16* __restrict__ dst, int C, int nx, int ny,
                       int nz, int64_t nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nxh = nx / 2, nyh = ny / 2, nzh = nz / 2;
  const int C8 = C / 8;
  const int64_t nc = (int64_t)nxh * nyh * nzh;
  if (i >= nb * nc * C8) return;
  const int c8 = (int)(i % C8);
  const int64_t q = (i / C8) % nc, bat = i / (C8 * nc);
  const int x = (int)(q % nxh), y = (int)((q / nxh) % nyh), z = (int)(q / ((int64_t)nxh * nyh));
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const bf16* sb = src + bat * (int64_t)nx * ny * nz * C;
  for (int d = 0; d < 8; d++) {
    const int X = 2 * x + (d & 1), Y = 2 * y + ((d >> 1) & 1), Z = 2 * z + (d >> 2);
    const bf16x8 v = *reinterpret_cast<const bf16x8*>(sb + (((int64_t)Z * ny + Y) * nx + X) * C + c8 * 8);
#pragma unroll
    for (int k = 0; k < 8; k++) s[k] += (float)v[k];
  }
  bf16x8 o;
#pragma unroll
  for (int k = 0; k < 8; k++) o[k] = (bf16)(s[k] * 0.125f);
  *reinterpret_cast<bf16x8*>(dst + (bat * nc + q) * C + c8 * 8) = o;
}

// dz = dy * (y > 0) (or dy when y is null), FP32 and bf16 copies; channels
// c >= C of the bf16 copy (row length Cp) are zero.
__global__ void k_relu_back(const float* __restrict__ dy, int ldy, const bf16* __restrict__ y,
                            int C, int Cp, float* __restrict__ dzf, bf16* __restrict__ dzb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * Cp) return;
  const int64_t p = i / Cp;
  const int c = (int)(i % Cp);
  float g = 0.f;
  if (c < C) {
    g = dy[p * ldy + c];
    if (y && !((float)y[p * C + c] > 0.f)) g = 0.f;
  }
  dzf[i] = g;
  dzb[i] = (bf16)g;
}

// split the gradient of a concatenation [up(src), skip] (row length C1+C2):
// dsrc[q][c] = sum of the 8 children (c < C1), dskip[p][c] = dcat[p][C1+c]
__global__ void k_split_cat(const float* __restrict__ dcat, int C1, int C2, float* __restrict__ dsrc,
                            float* __restrict__ dskip, int nx, int ny, int nz, int64_t nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nf = (int64_t)nx * ny * nz, nc = nf / 8;
  const int Ct = C1 + C2;
  const int64_t n1 = nb * nc * C1, n2 = nb * nf * C2;
  if (i < n1) {
    const int c = (int)(i % C1);
    const int64_t q = (i / C1) % nc, bat = i / (C1 * nc);
    const int nxh = nx / 2, nyh = ny / 2;
    const int x = (int)(q % nxh), y = (int)((q / nxh) % nyh), z = (int)(q / ((int64_t)nxh * nyh));
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
      const int X = 2 * x + (d & 1), Y = 2 * y + ((d >> 1) & 1), Z = 2 * z + (d >> 2);
      s += dcat[(bat * nf + ((int64_t)Z * ny + Y) * nx + X) * Ct + c];
    }
    dsrc[i] = s;
  } else if (i < n1 + n2) {
    const int64_t j = i - n1;
    const int c = (int)(j % C2);
    const int64_t p = j / C2;
    dskip[j] = dcat[p * Ct + C1 + c];
  }
}

// dfine[p][c] += dcoarse[p/2][c] / 8
__global__ void k_pool_back_add(const float* __restrict__ dc, float* __restrict__ df, int C, int nx,
                                int ny, int nz, int64_t nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nf = (int64_t)nx * ny * nz;
  if (i >= nb * nf * C) return;
  const int c = (int)(i % C);
  const int64_t p = (i / C) % nf, bat = i / (C * nf);
  const int x = (int)(p % nx), y = (int)((p / nx) % ny), z = (int)(p / ((int64_t)nx * ny));
  const int64_t q = ((int64_t)(z / 2) * (ny / 2) + y / 2) * (nx / 2) + x / 2;
  df[i] += 0.125f * dc[(bat * (nf / 8) + q) * C + c];
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (bf16)a[i];
}

// MSE: loss += sum (o-t)^2 / n; dout[p][0..7] = 2 (o-t)/n (channels 6, 7 zero)
__global__ void k_mse(const float* __restrict__ o, const float* __restrict__ t, float* __restrict__ dout,
                      int64_t nvox, double inv_n, double* loss) {
  __shared__ double red[UB];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0;
  if (i < nvox * 8) {
    const int64_t p = i / 8;
    const int c = (int)(i % 8);
    float g = 0.f;
    if (c < 6) {
      const float d = o[p * 6 + c] - t[p * 6 + c];
      s = (double)d * d * inv_n;
      g = (float)(2.0 * d * inv_n);
    }
    dout[i] = g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o2 = UB / 2; o2 > 0; o2 >>= 1) {
    if (threadIdx.x < o2) red[threadIdx.x] += red[threadIdx.x + o2];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int64_t n, float lr, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = 0.9f * m1[i] + 0.1f * gi;
  const float b = 0.999f * m2[i] + 0.001f * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
}

// ------------------------------------------------------------- features --
// sum of |v|^2 into per-block partials
__global__ void k_sumsq(const zd* __restrict__ v, int64_t n, double* part) {
  __shared__ double red[UB];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * UB + threadIdx.x; i < n; i += (int64_t)gridDim.x * UB)
    s += v[i].x * v[i].x + v[i].y * v[i].y;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = UB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// scale = [rms, 1/rms] with rms = ||v|| / sqrt(n) (1 for a zero vector)
__global__ void k_scale_fin(const double* part, int nb, int64_t n, double* scale) {
  __shared__ double red[UB];
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += UB) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = UB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double rms = sqrt(red[0] / (double)n);
    if (!(rms > 0)) rms = 1.0;
    scale[0] = rms;
    scale[1] = 1.0 / rms;
  }
}

// logc = log10(Re sa / Re sb) with sa recovered from a = 2 sqrt(Re sb)/(sa + conj sb)
__global__ void k_logc(const zd* __restrict__ asiga, const zd* __restrict__ csigb, const double* sq,
                       int64_t ncxy, int Nz, float* logc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncxy * Nz) return;
  const int z = (int)(i / ncxy);
  const zd a = asiga[i], sb = csigb[z];
  const zd sa = cdiv(mkz(2.0 * sq[z], 0.0), a) - mkz(sb.x, -sb.y);
  const double r = fmax(sa.x, 1e-12) / sb.x;
  logc[i] = (float)log10(r);
}

// x[p][8] from a residual r (slots s0,s1,s2 hold physical x,y,z); bf16 or FP32
__global__ void k_pack(const zd* __restrict__ r, int64_t ncell, int64_t ncxy, int s0, int s1, int s2,
                       const double* __restrict__ scale, const float* __restrict__ logc,
                       const float* __restrict__ logk, bf16* xb, float* xf) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncell) return;
  const double is = scale[1];
  const zd a = r[s0 * ncell + p], b = r[s1 * ncell + p], c = r[s2 * ncell + p];
  float v[8] = {(float)(a.x * is), (float)(a.y * is), (float)(b.x * is), (float)(b.y * is),
                (float)(c.x * is), (float)(c.y * is), logc[p], logk[p / ncxy]};
  if (xb) {
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = (bf16)v[k];
    *reinterpret_cast<bf16x8*>(xb + p * 8) = o;
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) xf[p * 8 + k] = v[k];
  }
}

// target t[p][6] = a(p) e_c(p) / rms
__global__ void k_target(const zd* __restrict__ e, const zd* __restrict__ asiga, int64_t ncell, int s0,
                         int s1, int s2, const double* __restrict__ scale, float* t) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncell) return;
  const double is = scale[1];
  const zd a = asiga[p];
  const zd ex = a * e[s0 * ncell + p], ey = a * e[s1 * ncell + p], ez = a * e[s2 * ncell + p];
  t[p * 6 + 0] = (float)(ex.x * is);
  t[p * 6 + 1] = (float)(ex.y * is);
  t[p * 6 + 2] = (float)(ey.x * is);
  t[p * 6 + 3] = (float)(ey.y * is);
  t[p * 6 + 4] = (float)(ez.x * is);
  t[p * 6 + 5] = (float)(ez.y * is);
}

// z_c(p) = rms * (o[p][2c] + i o[p][2c+1]) / a(p)
__global__ void k_unpack(const float* __restrict__ o, const zd* __restrict__ asiga, int64_t ncell,
                         int s0, int s1, int s2, const double* __restrict__ scale, zd* z) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncell) return;
  const double rms = scale[0];
  const zd ia = cdiv(mkz(rms, 0.0), asiga[p]);
  z[s0 * ncell + p] = ia * mkz(o[p * 6 + 0], o[p * 6 + 1]);
  z[s1 * ncell + p] = ia * mkz(o[p * 6 + 2], o[p * 6 + 3]);
  z[s2 * ncell + p] = ia * mkz(o[p * 6 + 4], o[p * 6 + 5]);
}

}  // namespace

// ------------------------------------------------------------------ host --
// (name, cin, cout, taps, relu): the architecture of the network
static const int kArch[11][4] = {{8, 16, 27, 1},  {16, 16, 27, 1}, {16, 32, 27, 1}, {32, 32, 27, 1},
                                 {32, 64, 27, 1}, {64, 64, 27, 1}, {96, 32, 27, 1}, {32, 32, 27, 1},
                                 {48, 16, 27, 1}, {16, 16, 27, 1}, {16, 6, 1, 0}};

int64_t unet_param_count() {
  int64_t c = 0;
  for (auto& a : kArch) c += (int64_t)a[1] * a[2] * a[0] + a[1];
  return c;
}

void unet_init(UNet& n, int) {
  n.L.clear();
  int64_t po = 0, wo = 0, wto = 0;
  for (auto& a : kArch) {
    UNetLayer l;
    l.cin = a[0];
    l.cout = a[1];
    l.ntap = a[2];
    l.relu = a[3];
    l.poff = po;
    po += (int64_t)l.cout * l.ntap * l.cin + l.cout;
    l.coutP = (int)up_to(l.cout, 16);
    l.kp = (int)up_to((int64_t)l.ntap * l.cin, 32);
    l.coutT = (int)up_to(l.cout, 8);
    l.cinP = (int)up_to(l.cin, 16);
    l.kpT = (int)up_to((int64_t)l.ntap * l.coutT, 32);
    l.woff = wo;
    l.wtoff = wto;
    wo += up_to((int64_t)l.coutP * l.kp, 128);
    wto += up_to((int64_t)l.cinP * l.kpT, 128);
    n.L.push_back(l);
  }
  n.nparams = po;
  n.nw = wo;
  n.nwt = wto;
  G2G_HIP(hipStreamCreate(&n.stream));
  G2G_HIP(hipMalloc(&n.params, sizeof(float) * po));
  G2G_HIP(hipMalloc(&n.grads, sizeof(float) * po));
  G2G_HIP(hipMalloc(&n.m1, sizeof(float) * po));
  G2G_HIP(hipMalloc(&n.m2, sizeof(float) * po));
  G2G_HIP(hipMalloc(&n.w, sizeof(bf16) * wo));
  G2G_HIP(hipMalloc(&n.wt, sizeof(bf16) * wto));
  G2G_HIP(hipMalloc(&n.dscal, sizeof(double) * 4));
  G2G_HIP(hipMemsetAsync(n.params, 0, sizeof(float) * po, n.stream));
  G2G_HIP(hipMemsetAsync(n.grads, 0, sizeof(float) * po, n.stream));
  G2G_HIP(hipMemsetAsync(n.m1, 0, sizeof(float) * po, n.stream));
  G2G_HIP(hipMemsetAsync(n.m2, 0, sizeof(float) * po, n.stream));
  G2G_HIP(hipStreamSynchronize(n.stream));
  n.adam_t = 0;
  n.wdirty = true;
}

void unet_free(UNet& n) {
  for (void* p : {(void*)n.params, (void*)n.grads, (void*)n.m1, (void*)n.m2, (void*)n.w, (void*)n.wt,
                  (void*)n.dscal, n.ws})
    if (p) (void)hipFree(p);
  if (n.stream) (void)hipStreamDestroy(n.stream);
  n = UNet();
}

void unet_reserve(UNet& n, int B, int nx, int ny, int nz, bool training) {
  G2G_CHECK(nx % 4 == 0 && ny % 4 == 0 && nz % 4 == 0 && nx > 0 && ny > 0 && nz > 0, -2,
            "U-Net volume dimensions must be positive multiples of 4");
  G2G_CHECK(B >= 1, -2, "batch must be >= 1");
  if (n.ws && n.B == B && n.nx == nx && n.ny == ny && n.nz == nz && (n.training || !training)) return;
  if (n.ws) {
    G2G_HIP(hipStreamSynchronize(n.stream));
    (void)hipFree(n.ws);
    n.ws = nullptr;
  }
  const int64_t n0 = (int64_t)B * nx * ny * nz, n1 = n0 / 8, n2 = n0 / 64;
  size_t off = 0;
  std::vector<std::pair<void**, size_t>> plan;
  auto add = [&](void** p, size_t bytes) {
    plan.push_back({p, off});
    off += (bytes + 255) / 256 * 256;
  };
  add((void**)&n.in, 2 * n0 * 8);
  const int64_t lvl[11] = {n0, n0, n1, n1, n2, n2, n1, n1, n0, n0, n0};
  for (int i = 0; i < 10; i++) add((void**)&n.act[i], 2 * lvl[i] * n.L[i].cout);
  n.act[10] = nullptr;
  add((void**)&n.p1, 2 * n1 * 16);
  add((void**)&n.p2, 2 * n2 * 32);
  add((void**)&n.out, 4 * n0 * 6);
  if (training) {
    add((void**)&n.dout, 4 * n0 * 8);
    add((void**)&n.f48, 4 * n0 * 48);
    add((void**)&n.fa, 4 * n0 * 16);
    add((void**)&n.fb, 4 * n0 * 4);
    add((void**)&n.de1, 4 * n0 * 16);
    add((void**)&n.de2, 4 * n1 * 32);
    add((void**)&n.dzf, 4 * n0 * 16);
    add((void**)&n.dzb, 2 * n0 * 16);
  } else {
    n.dout = n.f48 = n.fa = n.fb = n.de1 = n.de2 = n.dzf = nullptr;
    n.dzb = nullptr;
  }
  void* base = nullptr;
  G2G_HIP(hipMalloc(&base, off));
  for (auto& pr : plan) *pr.first = (char*)base + pr.second;
  n.ws = base;
  n.ws_bytes = off;
  n.B = B;
  n.nx = nx;
  n.ny = ny;
  n.nz = nz;
  n.training = training;
}

void unet_prepare_weights(UNet& n, hipStream_t s) {
  if (!n.wdirty) return;
  for (auto& l : n.L) {
    const int64_t tot = (int64_t)l.coutP * l.kp + (int64_t)l.cinP * l.kpT;
    hipLaunchKernelGGL(k_wprep, dim3(nblk(tot)), dim3(UB), 0, s, n.params + l.poff, l.cout, l.cin, l.ntap,
                       n.w + l.woff, l.coutP, l.kp, n.wt + l.wtoff, l.cinP, l.coutT, l.kpT);
  }
  G2G_HIP(hipGetLastError());
  n.wdirty = false;
}

static Src src(const bf16* p, int C, int up, int64_t bstride) {
  Src r;
  r.p = p;
  r.C = C;
  r.up = up;
  r.bstride = bstride;
  return r;
}

// y = conv(x) at level dims (nx, ny, nz), batch B
static void conv_fwd(UNet& n, int li, Src s1, Src s2, int nx, int ny, int nz, bf16* ybf, float* yf,
                     int ldy, hipStream_t s) {
  const UNetLayer& l = n.L[li];
  ConvArgs a;
  a.s1 = s1;
  a.s2 = s2;
  a.w = n.w + l.woff;
  a.kp = l.kp;
  a.ntap = l.ntap;
  a.cin = l.cin;
  a.bias = n.params + l.poff + (int64_t)l.cout * l.ntap * l.cin;
  a.cout = l.cout;
  a.relu = l.relu;
  a.ybf = ybf;
  a.yf = yf;
  a.ldy = ldy;
  a.ybstride = (int64_t)nx * ny * nz * ldy;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  G2G_CHECK(s1.C + s2.C == l.cin, -1, "U-Net layer input width mismatch");
  const int64_t nvox = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(k_conv, dim3(nblk(nvox, 256), l.coutP / 16, n.B), dim3(UB), 0, s, a);
}

// data gradient: dx[p][ci] (FP32, row length cin) from dzb[p][coutT] (bf16)
static void conv_dx(UNet& n, int li, const bf16* dzb, int nx, int ny, int nz, float* dx, hipStream_t s) {
  const UNetLayer& l = n.L[li];
  const int64_t nvox = (int64_t)nx * ny * nz;
  ConvArgs a;
  a.s1 = src(dzb, l.coutT, 0, nvox * l.coutT);
  a.s2 = src(nullptr, 0, 0, 0);
  a.w = n.wt + l.wtoff;
  a.kp = l.kpT;
  a.ntap = l.ntap;
  a.cin = l.coutT;
  a.bias = nullptr;
  a.cout = l.cin;
  a.relu = 0;
  a.ybf = nullptr;
  a.yf = dx;
  a.ldy = l.cin;
  a.ybstride = nvox * l.cin;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  hipLaunchKernelGGL(k_conv, dim3(nblk(nvox, 256), l.cinP / 16, n.B), dim3(UB), 0, s, a);
}

static void conv_dw(UNet& n, int li, const float* dzf, int ldz, Src s1, Src s2, int nx, int ny, int nz,
                    hipStream_t s) {
  const UNetLayer& l = n.L[li];
  DwArgs a;
  a.dz = dzf;
  a.ldz = ldz;
  a.cout = l.cout;
  a.cout4 = (int)up_to(l.cout, 4);
  a.s1 = s1;
  a.s2 = s2;
  a.cin = l.cin;
  a.ntap = l.ntap;
  a.dW = n.grads + l.poff;
  a.db = n.grads + l.poff + (int64_t)l.cout * l.ntap * l.cin;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  const int64_t nvox = (int64_t)nx * ny * nz;
  a.dzbstride = nvox * ldz;
  const int64_t nch = (nvox + DWV - 1) / DWV;
  a.chunks = (int)std::max<int64_t>(1, std::min<int64_t>(32, nch / 64));
  const size_t lds = sizeof(float) * DWV * (a.cout4 + l.cin);
  G2G_CHECK((a.cout4 / 4) * (l.cin / 4) <= UB, -1, "weight-gradient tile count exceeds the block");
  hipLaunchKernelGGL(k_dw, dim3((unsigned)((nch + a.chunks - 1) / a.chunks), l.ntap, n.B), dim3(UB), lds, s, a);
}

void unet_load_input(UNet& n, const float* x, hipStream_t s) {
  const int64_t cnt = (int64_t)n.B * n.nx * n.ny * n.nz * 8;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(cnt)), dim3(UB), 0, s, x, n.in, cnt);
}

void unet_forward(UNet& n, hipStream_t s) {
  G2G_CHECK(n.ws != nullptr, -4, "U-Net workspace not reserved");
  unet_prepare_weights(n, s);
  const int X0 = n.nx, Y0 = n.ny, Z0 = n.nz, X1 = X0 / 2, Y1 = Y0 / 2, Z1 = Z0 / 2, X2 = X1 / 2,
            Y2 = Y1 / 2, Z2 = Z1 / 2;
  const int64_t v0 = (int64_t)X0 * Y0 * Z0, v1 = v0 / 8, v2 = v0 / 64;
  const Src none = src(nullptr, 0, 0, 0);
  bf16** A = n.act;
  conv_fwd(n, 0, src(n.in, 8, 0, v0 * 8), none, X0, Y0, Z0, A[0], nullptr, 16, s);
  conv_fwd(n, 1, src(A[0], 16, 0, v0 * 16), none, X0, Y0, Z0, A[1], nullptr, 16, s);
  hipLaunchKernelGGL(k_pool, dim3(nblk((int64_t)n.B * v1 * 2)), dim3(UB), 0, s, A[1], n.p1, 16, X0, Y0, Z0,
                     (int64_t)n.B);
  conv_fwd(n, 2, src(n.p1, 16, 0, v1 * 16), none, X1, Y1, Z1, A[2], nullptr, 32, s);
  conv_fwd(n, 3, src(A[2], 32, 0, v1 * 32), none, X1, Y1, Z1, A[3], nullptr, 32, s);
  hipLaunchKernelGGL(k_pool, dim3(nblk((int64_t)n.B * v2 * 4)), dim3(UB), 0, s, A[3], n.p2, 32, X1, Y1, Z1,
                     (int64_t)n.B);
  conv_fwd(n, 4, src(n.p2, 32, 0, v2 * 32), none, X2, Y2, Z2, A[4], nullptr, 64, s);
  conv_fwd(n, 5, src(A[4], 64, 0, v2 * 64), none, X2, Y2, Z2, A[5], nullptr, 64, s);
  conv_fwd(n, 6, src(A[5], 64, 1, v2 * 64), src(A[3], 32, 0, v1 * 32), X1, Y1, Z1, A[6], nullptr, 32, s);
  conv_fwd(n, 7, src(A[6], 32, 0, v1 * 32), none, X1, Y1, Z1, A[7], nullptr, 32, s);
  conv_fwd(n, 8, src(A[7], 32, 1, v1 * 32), src(A[1], 16, 0, v0 * 16), X0, Y0, Z0, A[8], nullptr, 16, s);
  conv_fwd(n, 9, src(A[8], 16, 0, v0 * 16), none, X0, Y0, Z0, A[9], nullptr, 16, s);
  conv_fwd(n, 10, src(A[9], 16, 0, v0 * 16), none, X0, Y0, Z0, nullptr, n.out, 6, s);
  G2G_HIP(hipGetLastError());
}

// one layer of the backward pass: dz = dy*(y>0), weight gradient, data gradient into dx
static void layer_back(UNet& n, int li, const float* dy, int ldy, const bf16* y, Src s1, Src s2, int nx,
                       int ny, int nz, float* dx, hipStream_t s) {
  const UNetLayer& l = n.L[li];
  const int64_t nv = (int64_t)n.B * nx * ny * nz;
  hipLaunchKernelGGL(k_relu_back, dim3(nblk(nv * l.coutT)), dim3(UB), 0, s, dy, ldy, l.relu ? y : nullptr,
                     l.cout, l.coutT, n.dzf, n.dzb, nv);
  conv_dw(n, li, n.dzf, l.coutT, s1, s2, nx, ny, nz, s);
  if (dx) conv_dx(n, li, n.dzb, nx, ny, nz, dx, s);
}

void unet_backward(UNet& n, hipStream_t s) {
  G2G_CHECK(n.training, -4, "U-Net workspace not reserved for training");
  const int X0 = n.nx, Y0 = n.ny, Z0 = n.nz, X1 = X0 / 2, Y1 = Y0 / 2, Z1 = Z0 / 2, X2 = X1 / 2,
            Y2 = Y1 / 2, Z2 = Z1 / 2;
  const int64_t v0 = (int64_t)X0 * Y0 * Z0, v1 = v0 / 8, v2 = v0 / 64;
  const int64_t B = n.B;
  const Src none = src(nullptr, 0, 0, 0);
  bf16** A = n.act;
  G2G_HIP(hipMemsetAsync(n.grads, 0, sizeof(float) * n.nparams, s));
  layer_back(n, 10, n.dout, 8, nullptr, src(A[9], 16, 0, v0 * 16), none, X0, Y0, Z0, n.fa, s);
  layer_back(n, 9, n.fa, 16, A[9], src(A[8], 16, 0, v0 * 16), none, X0, Y0, Z0, n.fa, s);
  layer_back(n, 8, n.fa, 16, A[8], src(A[7], 32, 1, v1 * 32), src(A[1], 16, 0, v0 * 16), X0, Y0, Z0, n.f48,
             s);
  hipLaunchKernelGGL(k_split_cat, dim3(nblk(B * v1 * 32 + B * v0 * 16)), dim3(UB), 0, s, n.f48, 32, 16, n.fb,
                     n.de1, X0, Y0, Z0, B);
  layer_back(n, 7, n.fb, 32, A[7], src(A[6], 32, 0, v1 * 32), none, X1, Y1, Z1, n.fb, s);
  layer_back(n, 6, n.fb, 32, A[6], src(A[5], 64, 1, v2 * 64), src(A[3], 32, 0, v1 * 32), X1, Y1, Z1, n.f48,
             s);
  hipLaunchKernelGGL(k_split_cat, dim3(nblk(B * v2 * 64 + B * v1 * 32)), dim3(UB), 0, s, n.f48, 64, 32, n.fb,
                     n.de2, X1, Y1, Z1, B);
  layer_back(n, 5, n.fb, 64, A[5], src(A[4], 64, 0, v2 * 64), none, X2, Y2, Z2, n.fb, s);
  layer_back(n, 4, n.fb, 64, A[4], src(n.p2, 32, 0, v2 * 32), none, X2, Y2, Z2, n.fb, s);
  hipLaunchKernelGGL(k_pool_back_add, dim3(nblk(B * v1 * 32)), dim3(UB), 0, s, n.fb, n.de2, 32, X1, Y1, Z1, B);
  layer_back(n, 3, n.de2, 32, A[3], src(A[2], 32, 0, v1 * 32), none, X1, Y1, Z1, n.fb, s);
  layer_back(n, 2, n.fb, 32, A[2], src(n.p1, 16, 0, v1 * 16), none, X1, Y1, Z1, n.fb, s);
  hipLaunchKernelGGL(k_pool_back_add, dim3(nblk(B * v0 * 16)), dim3(UB), 0, s, n.fb, n.de1, 16, X0, Y0, Z0, B);
  layer_back(n, 1, n.de1, 16, A[1], src(A[0], 16, 0, v0 * 16), none, X0, Y0, Z0, n.fa, s);
  layer_back(n, 0, n.fa, 16, A[0], src(n.in, 8, 0, v0 * 8), none, X0, Y0, Z0, nullptr, s);
  G2G_HIP(hipGetLastError());
}

double unet_mse(UNet& n, const float* target, hipStream_t s) {
  const int64_t nv = (int64_t)n.B * n.nx * n.ny * n.nz;
  G2G_HIP(hipMemsetAsync(n.dscal, 0, sizeof(double), s));
  hipLaunchKernelGGL(k_mse, dim3(nblk(nv * 8)), dim3(UB), 0, s, n.out, target, n.dout, nv,
                     1.0 / (double)(nv * 6), n.dscal);
  G2G_HIP(hipGetLastError());
  double loss = 0;
  G2G_HIP(hipMemcpyAsync(&loss, n.dscal, sizeof(double), hipMemcpyDeviceToHost, s));
  G2G_HIP(hipStreamSynchronize(s));
  return loss;
}

void unet_adam(UNet& n, float lr, hipStream_t s) {
  n.adam_t++;
  const float c1 = 1.0f / (1.0f - powf(0.9f, (float)n.adam_t));
  const float c2 = 1.0f / (1.0f - powf(0.999f, (float)n.adam_t));
  hipLaunchKernelGGL(k_adam, dim3(nblk(n.nparams)), dim3(UB), 0, s, n.params, n.grads, n.m1, n.m2, n.nparams,
                     lr, c1, c2);
  G2G_HIP(hipGetLastError());
  n.wdirty = true;
}

// ------------------------------------------------------------ features --
void learned_prepare(LearnedCtx& c, Operator& op, UNet& net) {
  G2G_CHECK(op.have_sigma, -4, "conductivity not set");
  if (c.ncell != op.ncell()) learned_free(c);
  c.net = &net;
  c.op = &op;
  c.ncell = op.ncell();
  if (!c.logc) {
    G2G_HIP(hipMalloc(&c.logc, sizeof(float) * c.ncell));
    G2G_HIP(hipMalloc(&c.logk, sizeof(float) * op.Nz));
    G2G_HIP(hipMalloc(&c.scale, sizeof(double) * 2));
    G2G_HIP(hipMalloc(&c.part, sizeof(double) * 1024));
  }
  // frequency feature per layer: log10(|k_b| dx), |k_b|^2 = omega mu0 |sigma_b|
  std::vector<float> lk(op.Nz);
  for (int z = 0; z < op.Nz; z++) {
    const double kk = op.omega * (4e-7 * 3.14159265358979323846) * std::abs(std::complex<double>(op.csigb[z].x, op.csigb[z].y));
    lk[z] = (float)std::log10(std::sqrt(kk) * op.dx);
  }
  G2G_HIP(hipMemcpyAsync(c.logk, lk.data(), sizeof(float) * op.Nz, hipMemcpyHostToDevice, op.stream));
  zd* d_csigb = nullptr;
  G2G_HIP(hipMalloc(&d_csigb, sizeof(zd) * op.Nz));
  G2G_HIP(hipMemcpyAsync(d_csigb, op.csigb.data(), sizeof(zd) * op.Nz, hipMemcpyHostToDevice, op.stream));
  hipLaunchKernelGGL(k_logc, dim3(nblk(c.ncell)), dim3(UB), 0, op.stream, op.d64.asiga, d_csigb, op.d64.sq,
                     (int64_t)op.Nx * op.Ny, op.Nz, c.logc);
  G2G_HIP(hipGetLastError());
  G2G_HIP(hipStreamSynchronize(op.stream));
  (void)hipFree(d_csigb);
}

void learned_free(LearnedCtx& c) {
  for (void* p : {(void*)c.logc, (void*)c.logk, (void*)c.scale, (void*)c.part})
    if (p) (void)hipFree(p);
  c.logc = c.logk = nullptr;
  c.scale = c.part = nullptr;
  c.ncell = 0;
}

static void residual_scale(LearnedCtx& c, const zd* r, hipStream_t s) {
  const int64_t n = 3 * c.ncell;
  const int nb = (int)std::min<int64_t>(1024, (n + UB - 1) / UB);
  hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(UB), 0, s, r, n, c.part);
  hipLaunchKernelGGL(k_scale_fin, dim3(1), dim3(UB), 0, s, c.part, nb, n, c.scale);
}

void learned_features(LearnedCtx& c, const zd* r, const zd* e, float* x, float* t, hipStream_t s) {
  Operator& op = *c.op;
  residual_scale(c, r, s);
  const int* sl = op.comp_slot;
  hipLaunchKernelGGL(k_pack, dim3(nblk(c.ncell)), dim3(UB), 0, s, r, c.ncell, (int64_t)op.Nx * op.Ny, sl[0],
                     sl[1], sl[2], c.scale, c.logc, c.logk, (bf16*)nullptr, x);
  if (e && t)
    hipLaunchKernelGGL(k_target, dim3(nblk(c.ncell)), dim3(UB), 0, s, e, op.d64.asiga, c.ncell, sl[0], sl[1],
                       sl[2], c.scale, t);
  G2G_HIP(hipGetLastError());
}

void learned_apply(LearnedCtx& c, const zd* v, zd* z, hipStream_t s) {
  Operator& op = *c.op;
  UNet& n = *c.net;
  unet_reserve(n, 1, op.Nx, op.Ny, op.Nz, false);
  residual_scale(c, v, s);
  const int* sl = op.comp_slot;
  hipLaunchKernelGGL(k_pack, dim3(nblk(c.ncell)), dim3(UB), 0, s, v, c.ncell, (int64_t)op.Nx * op.Ny, sl[0],
                     sl[1], sl[2], c.scale, c.logc, c.logk, n.in, (float*)nullptr);
  unet_forward(n, s);
  hipLaunchKernelGGL(k_unpack, dim3(nblk(c.ncell)), dim3(UB), 0, s, n.out, op.d64.asiga, c.ncell, sl[0], sl[1],
                     sl[2], c.scale, z);
  G2G_HIP(hipGetLastError());
}

}  // namespace g2g
