// g2g_unet.hpp -- the learned right preconditioner: a 3-level 3D U-Net with
// BF16 weights and activations on MFMA (FP32 accumulation), its training
// kernels, and the hook that turns it into z = M(v) for the outer FGMRES.
// The reference has no counterpart: its right preconditioner is the inner
// GMRES of GFGMRES/seq_fgmres.F90:142-152.
#pragma once
#include <vector>

#include "g2g_common.hpp"

namespace g2g {

typedef __bf16 bf16;

struct UNetLayer {
  int cin, cout, ntap, relu;
  int64_t poff;     // offset of W[cout][ntap][cin] then b[cout] in the parameter vector
  int coutP, kp;    // forward weights: [coutP][kp], k = tap*cin + ci
  int coutT, cinP, kpT;  // transposed weights for the data gradient: [cinP][kpT], k = tap'*coutT + co
  int64_t woff, wtoff;   // element offsets in the packed bf16 weight arrays
};

struct UNet {
  std::vector<UNetLayer> L;
  int64_t nparams = 0, nw = 0, nwt = 0;
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
  bf16 *w = nullptr, *wt = nullptr;
  int adam_t = 0;
  bool wdirty = true;  // bf16 weights need regenerating from params
  hipStream_t stream = nullptr;
  // workspace for (B, nx, ny, nz); training adds the gradient buffers
  int B = 0, nx = 0, ny = 0, nz = 0;
  bool training = false;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  bf16 *in = nullptr, *act[11] = {}, *p1 = nullptr, *p2 = nullptr;
  float* out = nullptr;
  float *dout = nullptr, *f48 = nullptr, *fa = nullptr, *fb = nullptr, *de1 = nullptr,
        *de2 = nullptr, *dzf = nullptr;
  bf16* dzb = nullptr;
  double* dscal = nullptr;  // loss accumulator / residual scale
};

int64_t unet_param_count();
void unet_init(UNet& n, int seed_unused);
void unet_free(UNet& n);
void unet_reserve(UNet& n, int B, int nx, int ny, int nz, bool training);
void unet_prepare_weights(UNet& n, hipStream_t s);
// forward of the batch already packed in n.in; output FP32 [B][vox][6] in n.out
void unet_forward(UNet& n, hipStream_t s);
// gradient of sum(dy*out) for the last forward; dy FP32 [B][vox][8] in n.dout
void unet_backward(UNet& n, hipStream_t s);
// copy an FP32 [B][vox][8] input into n.in (bf16)
void unet_load_input(UNet& n, const float* x, hipStream_t s);
// MSE loss of n.out vs target [B][vox][6]; fills n.dout; returns the loss (syncs)
double unet_mse(UNet& n, const float* target, hipStream_t s);
void unet_adam(UNet& n, float lr, hipStream_t s);

struct Operator;
// Features of a residual r on the operator grid (complex128, memory slots as
// the operator's vectors): x [vox][8] = (r/rms re/im per physical component,
// log10(Re sa / Re sb), log10(|k_b| dx)); optional target t [vox][6] = a*e/rms.
struct LearnedCtx {
  UNet* net = nullptr;
  Operator* op = nullptr;
  float* logc = nullptr;    // per cell: log10(Re sa / Re sb)
  float* logk = nullptr;    // per layer: log10(|k_b| dx), |k_b|^2 = omega mu0 |sb|
  double* scale = nullptr;  // [rms, 1/rms]
  double* part = nullptr;   // reduction partials
  int64_t ncell = 0;
};
void learned_prepare(LearnedCtx& c, Operator& op, UNet& net);
void learned_free(LearnedCtx& c);
void learned_features(LearnedCtx& c, const zd* r, const zd* e, float* x, float* t, hipStream_t s);
// z = M(v): pack, network forward, unpack with 1/a and the residual scale
void learned_apply(LearnedCtx& c, const zd* v, zd* z, hipStream_t s);

}  // namespace g2g
