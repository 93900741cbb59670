// g2g_capi.hip -- extern "C" entry points (include/giem2g_amd.h).
#include <complex>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/giem2g_amd.h"
#include "g2g_op.hpp"
#include "g2g_solver.hpp"
#include "g2g_unet.hpp"

using namespace g2g;
using zc = std::complex<double>;

namespace {

thread_local std::string g_last_error;
void (*g_logger)(const char*) = nullptr;

constexpr double PI_D = 3.1415926535897932384626433832795028841971693993751;
constexpr double MU0 = 4.0 * PI_D * 1e-7;

void log_msg(const std::string& m) {
  if (g_logger) g_logger(m.c_str());
  else std::fprintf(stdout, "giem2g: %s\n", m.c_str());
}

template <typename F> int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code < 0 ? e.code : -1;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return -1;
  }
}

struct Handle {
  Operator op;
  int receivers = 0;  // receiver operators created on this operator (they hold a pointer to it)
  // the caller's conductivity array given to giem2g_set_anomaly_conductivity:
  // aliased as in the reference (giem2g_interfaces.F90:199), so every
  // giem2g_apply_ie_operator reads its current values
  const double* sigma_alias = nullptr;
  std::vector<double> hist;
  PrecondHook learned;
  UNet* net = nullptr;  // attached learned preconditioner (not owned)
  LearnedCtx lctx;
};

void learned_hook(void* ctx, const zd* v, zd* z) {
  LearnedCtx* c = static_cast<LearnedCtx*>(ctx);
  learned_apply(*c, v, z, c->op->stream);
}

UNet* NET(void* p) {
  G2G_CHECK(p != nullptr, -2, "null network handle");
  return static_cast<UNet*>(p);
}

void prepare_learned(Handle& h) {
  G2G_CHECK(h.net != nullptr, -4, "no learned preconditioner attached (giem2g_op_set_learned)");
  learned_prepare(h.lctx, h.op, *h.net);
}

Handle* H(void* p) {
  G2G_CHECK(p != nullptr, -2, "null operator handle");
  return static_cast<Handle*>(p);
}

struct RcHandle {
  RcOperator rc;
  Handle* owner = nullptr;
};

RcOperator& RC(void* p) {
  G2G_CHECK(p != nullptr, -2, "null receiver-operator handle");
  return static_cast<RcHandle*>(p)->rc;
}

// a host array through a temporary device buffer
struct DevCopy {
  zd* d = nullptr;
  DevCopy(const double* h, size_t n) {
    G2G_HIP(hipMalloc(&d, sizeof(zd) * n));
    if (h) G2G_HIP(hipMemcpy(d, h, sizeof(zd) * n, hipMemcpyHostToDevice));
  }
  ~DevCopy() {
    if (d) (void)hipFree(d);
  }
};

int get_layer(const Operator& op, double z) { return op_layer(op, z); }

void op_init(Operator& op, int Nx, int Ny, int Nz, double dx, double dy, const double* z) {
  // any even Nx, Ny (CalcLocalKernelSize, Integral_Equation_Module.F90:105-129,
  // rejects odd ones): powers of two run the radix-4 passes, other lengths
  // their factor plans
  G2G_CHECK(Nx >= 8 && Ny >= 8 && Nx % 2 == 0 && Ny % 2 == 0 && Nx <= 512 && Ny <= 512, -3,
            "Nx and Ny must be even and in [8,512]");
  G2G_CHECK(Nz >= 1 && Nz <= MAX_NZ, -3, "Nz must be in [1,256]");
  G2G_CHECK(dx > 0 && dy > 0, -3, "dx, dy must be positive");
  op.Nx = Nx;
  op.Ny = Ny;
  op.Nz = Nz;
  op.dx = dx;
  op.dy = dy;
  op.z.assign(z, z + Nz + 1);
  op.dz.resize(Nz);
  for (int i = 0; i < Nz; i++) {
    op.dz[i] = op.z[i + 1] - op.z[i];
    G2G_CHECK(op.dz[i] > 0, -3, "z borders must increase");
  }
  op.BS = 4 * (int64_t)Nz * (Nz + 1) / 2 + 2 * (int64_t)Nz * Nz;
  op.px = fft_plan(2 * Nx);
  op.py = fft_plan(2 * Ny);
  op.stream = nullptr;  // the default stream orders with torch's default stream
  for (auto& e : op.ev) G2G_HIP(hipEventCreate(&e));
  op_alloc_buffers<double>(op, op.d64, true);
  int dev = 0;
  G2G_HIP(hipGetDevice(&dev));
  op_noalloc_setup(dev);  // the allocation-free apply's scratch address on this device
}

void op_fini(Operator& op) {
  for (Arena& a : op.ws) a.release();
  if (op.adj) (void)hipFree(op.adj);
  op.adj = nullptr;
  if (op.csiga) (void)hipFree(op.csiga);
  op.csiga = nullptr;
  op.prof.release();
  op_free_buffers<double>(op.d64);
  op_free_buffers<float>(op.d32);
  op_free_tensor_scratch(op);
  for (auto& e : op.ev)
    if (e) (void)hipEventDestroy(e);
}

void set_background(Operator& op, int Nl, const double* csigb, const double* thick, double omega) {
  G2G_CHECK(Nl >= 1 && Nl <= 64, -3, "Nl must be in [1,64]");
  op.Nl = Nl;
  op.omega = omega;
  op.csigma.resize(Nl + 1);
  for (int i = 0; i <= Nl; i++) op.csigma[i] = mkz(csigb[2 * i], csigb[2 * i + 1]);
  op.thick.assign(Nl + 1, 0.0);
  for (int i = 1; i <= Nl - 1; i++) op.thick[i] = thick[i - 1];
  // Lnumber (mpi_saveload_module.F90:101-102)
  op.lnum.resize(op.Nz + 1);
  for (int i = 1; i <= op.Nz; i++) op.lnum[i] = get_layer(op, (op.z[i - 1] + op.z[i]) / 2.0);
  op.lnum[0] = get_layer(op, op.z[0] * (1.0 - 1e-7));
  op.csigb.resize(op.Nz);
  op.sqsigb.resize(op.Nz);
  for (int iz = 0; iz < op.Nz; iz++) {
    const int l = op.lnum[iz + 1];
    G2G_CHECK(l >= 1, -3, "anomaly layers must lie below the surface");
    op.csigb[iz] = op.csigma[l];
    op.sqsigb[iz] = std::sqrt(op.csigma[l].x);  // PreparePreconditioner (calc_ie_tensor_module.F90:303)
  }
  op.have_bkg = true;
  op.bkg_version++;
  op.have_tensor = false;
  op.tensor_own = false;
  op.have_g32 = false;
  // the per-cell scalings depend on the background: derive them again from
  // the conductivity last set
  if (op.have_sigma) op_set_sigma(op, op.csiga);
}

void build_tensor(Operator& op, int stage) {
  if (stage <= 0) {
    op_build_tensor(op);
    op.have_tensor = false;
    op.have_g32 = false;
    op.tensor_own = true;
  }
  if (stage != 0) {
    hipEvent_t a, b;
    G2G_HIP(hipEventCreate(&a));
    G2G_HIP(hipEventCreate(&b));
    G2G_HIP(hipEventRecord(a, op.stream));
    op_tensor_fft(op, op.d64.G);
    G2G_HIP(hipEventRecord(b, op.stream));
    G2G_HIP(hipEventSynchronize(b));
    float ms = 0;
    G2G_HIP(hipEventElapsedTime(&ms, a, b));
    op.cnt.t_tensor_fft = ms * 1e-3;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    op.have_tensor = true;
    if (op.d32.G) op_make_g32(op);
  }
}

void set_sigma_dev(Operator& op, const void* d) {
  op_set_sigma(op, static_cast<const zd*>(d));
}

void set_sigma_host(Operator& op, const double* h) {
  G2G_CHECK(h != nullptr, -2, "null conductivity");
  G2G_CHECK(op.have_bkg, -4, "background not set");
  const size_t n = op.ncell();
  if (!op.csiga) G2G_HIP(hipMalloc(&op.csiga, sizeof(zd) * n));
  G2G_HIP(hipStreamSynchronize(op.stream));
  G2G_HIP(hipMemcpy(op.csiga, h, sizeof(zd) * n, hipMemcpyHostToDevice));
  op_set_sigma(op, op.csiga);
}

// merge the event log of the profiling mode into the counters
void flush_prof(Operator& op) {
  if (!op.prof.marks.empty()) op.prof.flush(op.cnt);
}

void apply_host(Operator& op, const double* in, double* out) {
  const size_t n = op.nvec();
  zd *a, *b;
  G2G_HIP(hipMalloc(&a, sizeof(zd) * n));
  G2G_HIP(hipMalloc(&b, sizeof(zd) * n));
  G2G_HIP(hipMemcpy(a, in, sizeof(zd) * n, hipMemcpyHostToDevice));
  op_apply<double>(op, op.d64, a, b, op.stream, false);
  op.cnt.matvecs++;
  G2G_HIP(hipMemcpy(out, b, sizeof(zd) * n, hipMemcpyDeviceToHost));
  (void)hipFree(a);
  (void)hipFree(b);
}

// The normal (1-D) field of a unit plane wave: Calc_Apq at lambda = 0 with
// the TE column A(l,1,1) of the layer-coupling matrix (apq_module.F90:
// 230-266) and the surface normalisation t1 (sources_module.F90:36-42, 77-83).
struct PlaneWaveBasis {
  int Nl = 0;
  std::vector<zc> eta, q0, A1;
  std::vector<double> depth;
  zc t1;
};

PlaneWaveBasis plane_wave_basis(const Operator& op) {
  G2G_CHECK(op.have_bkg, -4, "background not set");
  const int Nl = op.Nl;
  std::vector<zc> cs(Nl + 1), k2(Nl + 1), eta(Nl + 1), el(Nl + 1);
  std::vector<double> depth(Nl + 1, 0.0);
  for (int i = 0; i <= Nl; i++) {
    cs[i] = zc(op.csigma[i].x, op.csigma[i].y);
    k2[i] = zc(0, 1) * cs[i] * MU0 * op.omega;
  }
  for (int i = 1; i <= Nl - 1; i++) depth[i] = depth[i - 1] + op.thick[i];
  const double lm = 0.0;
  for (int i = 1; i <= Nl; i++) eta[i] = std::sqrt(lm * lm - k2[i]);
  for (int i = 1; i <= Nl - 1; i++) el[i] = std::exp(-2.0 * eta[i] * op.thick[i]);
  el[Nl] = 0;
  std::vector<std::array<zc, 2>> p(Nl + 1), q(Nl + 1), Arr(Nl + 1), w(Nl + 1);
  {
    const zc eta0 = std::sqrt(lm * lm - k2[0]), ceta = eta0 / eta[1];
    p[1][0] = (1.0 - ceta) / (1.0 + ceta);
    const zc cg = cs[0] / cs[1];
    p[1][1] = (cg - ceta) / (cg + ceta);
    for (int i = 2; i <= Nl; i++) {
      zc a[2];
      a[0] = eta[i - 1] / eta[i];
      a[1] = a[0] * cs[i] / cs[i - 1];
      for (int k = 0; k < 2; k++) {
        zc ww = el[i - 1] * p[i - 1][k];
        ww = (ww - 1.0) / (ww + 1.0);
        ww = ww * a[k];
        p[i][k] = (1.0 + ww) / (1.0 - ww);
      }
    }
  }
  q[Nl][0] = q[Nl][1] = 0;
  if (Nl > 1) {
    zc a[2];
    a[0] = eta[Nl] / eta[Nl - 1];
    a[1] = a[0] * cs[Nl - 1] / cs[Nl];
    for (int k = 0; k < 2; k++) q[Nl - 1][k] = (1.0 - a[k]) / (1.0 + a[k]);
    for (int i = Nl - 2; i >= 1; i--) {
      a[0] = eta[i + 1] / eta[i];
      a[1] = a[0] * cs[i] / cs[i + 1];
      for (int k = 0; k < 2; k++) {
        zc ww = el[i + 1] * q[i + 1][k];
        ww = (ww - 1.0) / (ww + 1.0);
        ww = ww * a[k];
        q[i][k] = (1.0 + ww) / (1.0 - ww);
      }
    }
  }
  for (int i = 1; i <= Nl; i++)
    for (int k = 0; k < 2; k++) Arr[i][k] = 1.0 / eta[i] / (1.0 - p[i][k] * q[i][k] * el[i]);
  // A(I,1,1) column of the layer-coupling matrix, TE part only
  for (int i = 1; i <= Nl - 1; i++) {
    zc tmp = (k2[i] - k2[i + 1]);
    tmp = depth[i] * tmp;
    tmp = tmp / (eta[i + 1] + eta[i]);
    for (int k = 0; k < 2; k++) w[i][k] = (1.0 + p[i + 1][k]) / (1.0 + p[i][k] * el[i]) * std::exp(tmp);
  }
  std::vector<zc> A1(Nl + 1);  // A(l,1,1) = A(1,l,1)
  for (int l = 1; l <= Nl; l++) {
    zc a = Arr[l][0];
    for (int j = l - 1; j >= 1; j--) a = a * w[j][0];
    A1[l] = a;
  }
  zc t1 = Nl > 1 ? 1.0 + q[1][0] * std::exp(-depth[1] * eta[1] * 2.0) : zc(1.0);
  t1 = A1[1] * t1;
  PlaneWaveBasis b;
  b.Nl = Nl;
  b.eta = eta;
  b.A1 = A1;
  b.depth = depth;
  b.t1 = t1;
  b.q0.resize(Nl + 1);
  for (int i = 1; i <= Nl; i++) b.q0[i] = q[i][0];
  return b;
}

// PlaneWaveIntegral (sources_module.F90:67-101): cell averages over the anomaly.
void plane_wave(const Operator& op, int pol, double* Eh) {
  G2G_CHECK(pol == 0 || pol == 1, -3, "pol must be 0 (EX) or 1 (EY)");
  const PlaneWaveBasis b = plane_wave_basis(op);
  const int Nl = b.Nl;
  const std::vector<zc>& eta = b.eta;
  const std::vector<double>& depth = b.depth;
  const int64_t nxy = (int64_t)op.Nx * op.Ny, nc = op.ncell();
  std::memset(Eh, 0, sizeof(double) * 2 * 3 * nc);
  const int slot = op.comp_slot[pol];
  for (int Iz = 1; Iz <= op.Nz; Iz++) {
    const double dz = op.z[Iz] - op.z[Iz - 1];
    const int l = op.lnum[Iz];
    const zc e1 = std::exp(-eta[l] * op.z[Iz - 1]), e2 = std::exp(-eta[l] * op.z[Iz]);
    zc e3 = 0, e4 = 0;
    if (l < Nl) {
      e3 = b.q0[l] * std::exp(-eta[l] * (2.0 * depth[l] - op.z[Iz - 1]));
      e4 = b.q0[l] * std::exp(-eta[l] * (2.0 * depth[l] - op.z[Iz]));
    }
    zc G = (e1 - e2 + e4 - e3) * b.A1[l] / b.t1;
    G = G / dz / eta[l];
    double* dst = Eh + 2 * (slot * nc + (int64_t)(Iz - 1) * nxy);
    for (int64_t i = 0; i < nxy; i++) {
      dst[2 * i] = G.real();
      dst[2 * i + 1] = G.imag();
    }
  }
}

// PlaneWave (sources_module.F90:26-66): the field at receiver depths, F0
// complex (Nr,6) with components Ex,Ey,Ez,Hx,Hy,Hz.
void plane_wave_at(const Operator& op, int pol, int Nr, const double* zr, double* F0) {
  G2G_CHECK(pol == 0 || pol == 1, -3, "pol must be 0 (EX) or 1 (EY)");
  const PlaneWaveBasis b = plane_wave_basis(op);
  const zc iwm(0.0, MU0 * op.omega);
  std::memset(F0, 0, sizeof(double) * 2 * 6 * Nr);
  for (int r = 0; r < Nr; r++) {
    const int l = std::max(get_layer(op, zr[r]), 1);
    const zc t2 = b.A1[l] / b.t1;
    const zc e1 = std::exp(-b.eta[l] * zr[r]);
    const zc e2 = l < b.Nl ? b.q0[l] * std::exp(-b.eta[l] * (2.0 * b.depth[l] - zr[r])) : zc(0.0);
    const zc G0 = t2 * (e1 + e2);
    const zc G1 = -t2 * (e1 - e2) * b.eta[l] / iwm;
    const int ce = pol == 0 ? 0 : 1, ch = pol == 0 ? 4 : 3;
    const zc gh = pol == 0 ? G1 : -G1;
    F0[2 * (r + Nr * ce)] = G0.real();
    F0[2 * (r + Nr * ce) + 1] = G0.imag();
    F0[2 * (r + Nr * ch)] = gh.real();
    F0[2 * (r + Nr * ch) + 1] = gh.imag();
  }
}

SolveParams params_from(const giem2g_solver_params* p) {
  SolveParams s;
  if (p) {
    s.outer_m = p->outer_m;
    s.inner_m = p->inner_m;
    s.max_restart = p->max_restart;
    s.tol = p->tol;
    s.inner_precision = p->inner_precision;
    s.precond = p->precond;
    s.learned_patience = p->learned_patience == 0 ? 2 : (p->learned_patience < 0 ? 0 : p->learned_patience);
    s.inner_tol = p->inner_tol;
  }
  G2G_CHECK(s.outer_m >= 1 && s.outer_m <= 63 && s.inner_m >= 1 && s.inner_m <= 63, -3,
            "buffer sizes must be in [1,63]");
  G2G_CHECK(s.max_restart >= 1 && s.tol > 0, -3, "bad max_restart or tol");
  G2G_CHECK(s.inner_tol >= -1e6 && s.inner_tol < 1, -3, "inner_tol must be below 1 (negative: relaxed)");
  return s;
}

void solve_dev(Handle& h, const zd* Ebkg, const zd* E0, const giem2g_solver_params* pp,
               zd* Esol, giem2g_solve_info* info) {
  Operator& op = h.op;
  G2G_CHECK(op.have_tensor && op.have_sigma, -4, "operator not ready");
  const SolveParams p = params_from(pp);
  const int64_t N = op.nvec();
  zd* rhs = (zd*)op.ws[0].get(WS_RHS, sizeof(zd) * N);
  zd* guess = (zd*)op.ws[0].get(WS_GUESS, sizeof(zd) * N);
  hipEvent_t a, b;
  G2G_HIP(hipEventCreate(&a));
  G2G_HIP(hipEventCreate(&b));
  const int64_t m0 = op.cnt.matvecs, m32 = op.cnt.matvecs32;
  h.hist.clear();
  SolveResult r;
  try {
    G2G_HIP(hipEventRecord(a, op.stream));
    solve_scale(op, rhs, Ebkg, 0);  // rhs = sqrt(Re sb) E_bkg (ie_solver_module.F90:64-66)
    if (E0) solve_scale(op, guess, E0, 1);
    else G2G_HIP(hipMemcpyAsync(guess, rhs, sizeof(zd) * N, hipMemcpyDeviceToDevice, op.stream));
    if (p.precond != 0) prepare_learned(h);
    r = nested_solve(op, rhs, guess, Esol, p, &h.hist, h.learned.apply ? &h.learned : nullptr);
    solve_scale(op, Esol, Esol, 2);  // E = asiga x (:85-97)
    G2G_HIP(hipEventRecord(b, op.stream));
    G2G_HIP(hipEventSynchronize(b));
    flush_prof(op);
  } catch (...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  float ms = 0;
  G2G_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (info) {
    info->outer_iterations = r.iterations;
    info->status = r.stat;
    info->arnoldi_residual = r.bea;
    info->true_residual = r.be;
    info->matvecs_fp64 = op.cnt.matvecs - m0;
    info->matvecs_fp32 = op.cnt.matvecs32 - m32;
    info->precond_calls = r.precond_calls;
    info->precond_fallbacks = r.precond_fallbacks;
    info->seconds = ms * 1e-3;
  }
  if (pp && pp->verbose) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "GFGMRES status %d after %d outer iterations, Arnoldi %.3e, true %.3e",
                  r.stat, r.iterations, r.bea, r.be);
    log_msg(buf);
  }
}

}  // namespace

// ====================================================================== //
extern "C" {

const char* giem2g_last_error(void) { return g_last_error.c_str(); }

int giem2g_op_create(int Nx, int Ny, int Nz, double dx, double dy, const double* z, void** out) {
  return guard([&] {
    G2G_CHECK(out && z, -2, "null argument");
    auto h = std::make_unique<Handle>();
    op_init(h->op, Nx, Ny, Nz, dx, dy, z);
    *out = h.release();
  });
}

int giem2g_op_destroy(void* p) {
  return guard([&] {
    Handle* h = H(p);
    G2G_CHECK(h->receivers == 0, -4, "receiver operators still use this operator (giem2g_rc_destroy them first)");
    learned_free(h->lctx);
    op_fini(h->op);
    delete h;
  });
}

int giem2g_set_component_order(void* p, int order) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_CHECK(order == GIEM2G_ORDER_XYZ || order == GIEM2G_ORDER_YXZ, -3, "bad component order");
    op.comp_slot[0] = order == GIEM2G_ORDER_XYZ ? 0 : 1;
    op.comp_slot[1] = order == GIEM2G_ORDER_XYZ ? 1 : 0;
    op.comp_slot[2] = 2;
  });
}

int giem2g_op_set_background(void* p, int Nl, const double* csigb, const double* thick, double omega) {
  return guard([&] { set_background(H(p)->op, Nl, csigb, thick, omega); });
}

int giem2g_op_build_tensor(void* p) {
  return guard([&] { build_tensor(H(p)->op, -1); });
}

int giem2g_op_tensor_stage(void* p, int kind) {
  return guard([&] {
    G2G_CHECK(kind == 0 || kind == 1, -3, "kind must be 0 or 1");
    build_tensor(H(p)->op, kind);
  });
}

int giem2g_op_get_tensor(void* p, double* G) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_HIP(hipMemcpy(G, op.d64.G, sizeof(zd) * op.Nx * op.Ny * op.BS, hipMemcpyDeviceToHost));
  });
}

int giem2g_op_put_tensor(void* p, const double* G, int spectral) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_HIP(hipMemcpy(op.d64.G, G, sizeof(zd) * op.Nx * op.Ny * op.BS, hipMemcpyHostToDevice));
    op.have_tensor = spectral != 0;
    op.tensor_own = false;
    op.have_g32 = false;
    if (op.have_tensor && op.d32.G) op_make_g32(op);
  });
}

int giem2g_op_set_sigma(void* p, const double* csiga) {
  return guard([&] {
    Handle* h = H(p);
    set_sigma_host(h->op, csiga);
    h->sigma_alias = nullptr;
  });
}

int giem2g_op_set_sigma_dev(void* p, const void* csiga) {
  return guard([&] {
    Handle* h = H(p);
    G2G_CHECK(csiga != nullptr, -2, "null conductivity");
    set_sigma_dev(h->op, csiga);
    h->sigma_alias = nullptr;
  });
}

int giem2g_op_set_geometry(void* p, double dx, double dy, const double* z) {
  return guard([&] {
    G2G_CHECK(z != nullptr, -2, "null layer borders");
    Operator& op = H(p)->op;
    op_set_geometry(op, dx, dy, z);
    if (op.have_bkg) {
      // Lnumber, csigb and the scalings follow the new layer borders; the tensor must be built again
      std::vector<double> cs(2 * (op.Nl + 1)), th(op.Nl > 1 ? op.Nl - 1 : 1, 0.0);
      for (int i = 0; i <= op.Nl; i++) {
        cs[2 * i] = op.csigma[i].x;
        cs[2 * i + 1] = op.csigma[i].y;
      }
      for (int i = 1; i <= op.Nl - 1; i++) th[i - 1] = op.thick[i];
      set_background(op, op.Nl, cs.data(), th.data(), op.omega);
    }
  });
}

int giem2g_op_set_profiling(void* p, int on) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_HIP(hipStreamSynchronize(op.stream));
    flush_prof(op);
    op.prof.on = on != 0;
  });
}

int giem2g_op_counters(void* p, double* o, int n) {
  return guard([&] {
    G2G_CHECK(o != nullptr && n >= 0, -2, "bad output array");
    Operator& op = H(p)->op;
    flush_prof(op);
    const Counters& c = op.cnt;
    const double v[20] = {(double)c.matvecs, (double)c.matvecs32, (double)c.dotprods, c.t_apply,
                          c.t_fwd, c.t_contr, c.t_bwd, c.t_dot, c.t_tensor, c.t_tensor_fft,
                          (double)(op.d64.bytes + op.d32.bytes), (double)c.gs_passes, c.krylov_bytes,
                          (double)c.contr_launches[0], (double)c.contr_rhs[0], c.t_contr_p[0],
                          (double)c.contr_launches[1], (double)c.contr_rhs[1], c.t_contr_p[1],
                          (double)(op.prof.on ? 1 : 0)};
    for (int i = 0; i < n && i < 20; i++) o[i] = v[i];
  });
}

int giem2g_op_reset_counters(void* p) {
  return guard([&] {
    Operator& op = H(p)->op;
    flush_prof(op);
    op.cnt = Counters();
  });
}

int giem2g_op_apply_dev(void* p, const void* in, void* out) {
  return guard([&] {
    Operator& op = H(p)->op;
    op_apply<double>(op, op.d64, static_cast<const zd*>(in), static_cast<zd*>(out), op.stream,
                     false);
    op.cnt.matvecs++;
  });
}

int giem2g_op_apply_dev_f32(void* p, const void* in, void* out) {
  return guard([&] {
    Operator& op = H(p)->op;
    if (!op.have_g32) op_make_g32(op);
    op_apply<float>(op, op.d32, static_cast<const zf*>(in), static_cast<zf*>(out), op.stream,
                    false);
    op.cnt.matvecs32++;
  });
}

int giem2g_op_apply_pair_dev(void* p, int precision, const void* in0, const void* in1, void* out0,
                             void* out1) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_CHECK(precision == 0 || precision == 1, -3, "precision must be 0 or 1");
    if (precision == 0) {
      op_apply2<double>(op, op.d64, static_cast<const zd*>(in0), static_cast<const zd*>(in1),
                        static_cast<zd*>(out0), static_cast<zd*>(out1), op.stream);
      op.cnt.matvecs += 2;
    } else {
      if (!op.have_g32) op_make_g32(op);
      op_apply2<float>(op, op.d32, static_cast<const zf*>(in0), static_cast<const zf*>(in1),
                       static_cast<zf*>(out0), static_cast<zf*>(out1), op.stream);
      op.cnt.matvecs32 += 2;
    }
  });
}

int giem2g_op_apply_batch_dev(void* p, int precision, int nrhs, const void* in, void* out) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_CHECK(precision == 0 || precision == 1, -3, "precision must be 0 or 1");
    G2G_CHECK(nrhs >= 1, -3, "nrhs must be at least 1");
    if (precision == 0) {
      op_apply_batch<double>(op, op.d64, nrhs, static_cast<const zd*>(in), static_cast<zd*>(out), op.stream);
      op.cnt.matvecs += nrhs;
    } else {
      if (!op.have_g32) op_make_g32(op);
      op_apply_batch<float>(op, op.d32, nrhs, static_cast<const zf*>(in), static_cast<zf*>(out), op.stream);
      op.cnt.matvecs32 += nrhs;
    }
  });
}

int giem2g_op_apply_batch_dev_noalloc(void* p, int precision, int nrhs, const void* in, void* out) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_CHECK(precision == 0 || precision == 1, -3, "precision must be 0 or 1");
    G2G_CHECK(nrhs >= 1, -3, "nrhs must be at least 1");
    if (precision == 0) {
      op_apply_batch_noalloc<double>(op, op.d64, nrhs, static_cast<const zd*>(in), static_cast<zd*>(out),
                                     op.stream);
      op.cnt.matvecs += nrhs;
    } else {
      G2G_CHECK(op.have_g32, -4,
                "complex64 operator not made yet: make it with a complex64 apply or batch before the "
                "allocation-free apply");
      op_apply_batch_noalloc<float>(op, op.d32, nrhs, static_cast<const zf*>(in), static_cast<zf*>(out),
                                    op.stream);
      op.cnt.matvecs32 += nrhs;
    }
  });
}

int giem2g_op_apply_host(void* p, const double* in, double* out) {
  return guard([&] { apply_host(H(p)->op, in, out); });
}

int giem2g_op_stream(void* p, void** s) {
  return guard([&] { *s = (void*)H(p)->op.stream; });
}

int giem2g_op_sync(void* p) {
  return guard([&] { G2G_HIP(hipStreamSynchronize(H(p)->op.stream)); });
}

int giem2g_op_stats(void* p, double* o) {
  return guard([&] {
    Operator& op = H(p)->op;
    flush_prof(op);
    const Counters& c = op.cnt;
    const double v[11] = {(double)c.matvecs, (double)c.matvecs32, (double)c.dotprods, c.t_apply,
                          c.t_fwd, c.t_contr, c.t_bwd, c.t_dot, c.t_tensor, c.t_tensor_fft,
                          (double)(op.d64.bytes + op.d32.bytes)};
    std::memcpy(o, v, sizeof v);
  });
}

int giem2g_op_time_apply(void* p, int precision, const void* in, void* out, int reps, double* ms3) {
  return guard([&] {
    Operator& op = H(p)->op;
    G2G_CHECK(reps >= 1, -3, "reps must be >= 1");
    if (precision == 1 && !op.have_g32) op_make_g32(op);
    double acc[3] = {0, 0, 0};
    for (int r = 0; r < reps; r++) {
      if (precision == 0)
        op_apply<double>(op, op.d64, (const zd*)in, (zd*)out, op.stream, true);
      else
        op_apply<float>(op, op.d32, (const zf*)in, (zf*)out, op.stream, true);
      G2G_HIP(hipEventSynchronize(op.ev[3]));
      float t[3];
      for (int k = 0; k < 3; k++) G2G_HIP(hipEventElapsedTime(&t[k], op.ev[k], op.ev[k + 1]));
      for (int k = 0; k < 3; k++) acc[k] += t[k];
    }
    for (int k = 0; k < 3; k++) ms3[k] = acc[k] / reps;
    op.cnt.t_fwd += acc[0] * 1e-3;
    op.cnt.t_contr += acc[1] * 1e-3;
    op.cnt.t_bwd += acc[2] * 1e-3;
    op.cnt.t_apply += (acc[0] + acc[1] + acc[2]) * 1e-3;
    if (precision == 0) op.cnt.matvecs += reps;
    else op.cnt.matvecs32 += reps;
  });
}

int giem2g_plane_wave(void* p, int pol, double* E) {
  return guard([&] { plane_wave(H(p)->op, pol, E); });
}

int giem2g_solve_dev(void* p, const void* Ebkg, const void* E0, const giem2g_solver_params* prm,
                     void* Esol, giem2g_solve_info* info) {
  return guard([&] {
    solve_dev(*H(p), (const zd*)Ebkg, (const zd*)E0, prm, (zd*)Esol, info);
  });
}

int giem2g_solve_pair_dev(void* p, const void* Ebkg0, const void* Ebkg1, const void* E00, const void* E01,
                          const giem2g_solver_params* pp, void* Esol0, void* Esol1, giem2g_solve_info* info2,
                          long long* paired) {
  return guard([&] {
    Handle& h = *H(p);
    Operator& op = h.op;
    G2G_CHECK(op.have_tensor && op.have_sigma, -4, "operator not ready");
    G2G_CHECK(Ebkg0 && Ebkg1 && Esol0 && Esol1, -2, "null vector");
    const SolveParams prm = params_from(pp);
    const int64_t N = op.nvec();
    zd* work = (zd*)op.ws[0].get(WS_PAIR, sizeof(zd) * N * 4);
    zd* rhs[2] = {work, work + N};
    zd* guess[2] = {work + 2 * N, work + 3 * N};
    const zd* eb[2] = {(const zd*)Ebkg0, (const zd*)Ebkg1};
    const zd* e0[2] = {(const zd*)E00, (const zd*)E01};
    zd* es[2] = {(zd*)Esol0, (zd*)Esol1};
    hipEvent_t a, b;
    G2G_HIP(hipEventCreate(&a));
    G2G_HIP(hipEventCreate(&b));
    PairResult r;
    try {
      G2G_HIP(hipEventRecord(a, op.stream));
      for (int k = 0; k < 2; k++) {
        solve_scale(op, rhs[k], eb[k], 0);
        if (e0[k]) solve_scale(op, guess[k], e0[k], 1);
        else G2G_HIP(hipMemcpyAsync(guess[k], rhs[k], sizeof(zd) * N, hipMemcpyDeviceToDevice, op.stream));
      }
      const zd* const cb[2] = {rhs[0], rhs[1]};
      const zd* const cg[2] = {guess[0], guess[1]};
      r = nested_solve_pair(op, cb, cg, es, prm);
      for (int k = 0; k < 2; k++) solve_scale(op, es[k], es[k], 2);
      G2G_HIP(hipEventRecord(b, op.stream));
      G2G_HIP(hipEventSynchronize(b));
      flush_prof(op);
    } catch (...) {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      throw;
    }
    float ms = 0;
    G2G_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (info2) {
      for (int k = 0; k < 2; k++) {
        giem2g_solve_info& in = info2[k];
        in.outer_iterations = r.r[k].iterations;
        in.status = r.r[k].stat;
        in.arnoldi_residual = r.r[k].bea;
        in.true_residual = r.r[k].be;
        in.matvecs_fp64 = r.mv64[k];
        in.matvecs_fp32 = r.mv32[k];
        in.precond_calls = 0;
        in.precond_fallbacks = 0;
        in.seconds = ms * 1e-3;
      }
    }
    if (paired) {
      paired[0] = r.paired;
      paired[1] = r.alone;
    }
  });
}

int giem2g_solve(void* p, const double* Ebkg, const double* E0, const giem2g_solver_params* prm,
                 double* Esol, giem2g_solve_info* info) {
  return guard([&] {
    Handle* h = H(p);
    const size_t n = h->op.nvec();
    zd *a = nullptr, *b = nullptr, *c = nullptr;
    G2G_HIP(hipMalloc(&a, sizeof(zd) * n));
    G2G_HIP(hipMalloc(&c, sizeof(zd) * n));
    if (E0) G2G_HIP(hipMalloc(&b, sizeof(zd) * n));
    G2G_HIP(hipMemcpy(a, Ebkg, sizeof(zd) * n, hipMemcpyHostToDevice));
    if (E0) G2G_HIP(hipMemcpy(b, E0, sizeof(zd) * n, hipMemcpyHostToDevice));
    try {
      solve_dev(*h, a, b, prm, c, info);
      G2G_HIP(hipMemcpy(Esol, c, sizeof(zd) * n, hipMemcpyDeviceToHost));
    } catch (...) {
      (void)hipFree(a); (void)hipFree(b); (void)hipFree(c);
      throw;
    }
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(c);
  });
}

int giem2g_last_history(void* p, double* out, int maxlen, int* len) {
  return guard([&] {
    Handle* h = H(p);
    const int n = (int)h->hist.size();
    for (int i = 0; i < n && i < maxlen; i++) out[i] = h->hist[i];
    *len = n;
  });
}

int giem2g_op_set_contraction(void* p, int mode) {
  return guard([&] {
    Handle* h = H(p);
    G2G_CHECK(mode >= 0 && mode <= 6, -3, "contraction mode must be 0 to 6");
    G2G_CHECK(mode != 2 || op_pair_supported(h->op), -3,
              "the matrix-core contraction needs Nz from 8 to 64");
    G2G_CHECK((mode != 4 && mode != 5) || (h->op.Nz >= 8 && h->op.Nz <= 32), -3,
              "the small-block matrix-core contraction needs Nz from 8 to 32");
    G2G_CHECK((mode != 5 && mode != 6) || op_mirror_ok(h->op), -3,
              "the x-y mirror needs Nx = Ny, dx = dy and a tensor built by the library");
    G2G_CHECK(mode != 6 || op_pair_supported(h->op), -3,
              "the matrix-core contraction needs Nz from 8 to 64");
    h->op.contraction = mode;
  });
}

int giem2g_op_scale(void* p, int mode, const void* in_dev, void* out_dev) {
  return guard([&] {
    Handle* h = H(p);
    G2G_CHECK(h->op.have_sigma, -4, "conductivity not set");
    G2G_CHECK(mode >= 0 && mode <= 2, -3, "mode must be 0, 1 or 2");
    solve_scale(h->op, (zd*)out_dev, (const zd*)in_dev, mode);
    G2G_HIP(hipStreamSynchronize(h->op.stream));
  });
}

// ------------------------------------------------ learned preconditioner ---
int giem2g_unet_create(void** net) {
  return guard([&] {
    G2G_CHECK(net, -2, "null output pointer");
    std::unique_ptr<UNet> n(new UNet());
    unet_init(*n, 0);
    *net = n.release();
  });
}

int giem2g_unet_destroy(void* net) {
  return guard([&] {
    UNet* n = NET(net);
    unet_free(*n);
    delete n;
  });
}

int giem2g_unet_nparams(long long* n) {
  return guard([&] {
    G2G_CHECK(n, -2, "null output pointer");
    *n = unet_param_count();
  });
}

int giem2g_unet_set_params(void* net, const float* host) {
  return guard([&] {
    UNet* n = NET(net);
    G2G_HIP(hipMemcpy(n->params, host, sizeof(float) * n->nparams, hipMemcpyHostToDevice));
    G2G_HIP(hipMemset(n->m1, 0, sizeof(float) * n->nparams));
    G2G_HIP(hipMemset(n->m2, 0, sizeof(float) * n->nparams));
    n->adam_t = 0;
    n->wdirty = true;
  });
}

int giem2g_unet_get_params(void* net, float* host) {
  return guard([&] {
    UNet* n = NET(net);
    G2G_HIP(hipStreamSynchronize(n->stream));
    G2G_HIP(hipMemcpy(host, n->params, sizeof(float) * n->nparams, hipMemcpyDeviceToHost));
  });
}

int giem2g_unet_get_grads(void* net, float* host) {
  return guard([&] {
    UNet* n = NET(net);
    G2G_HIP(hipStreamSynchronize(n->stream));
    G2G_HIP(hipMemcpy(host, n->grads, sizeof(float) * n->nparams, hipMemcpyDeviceToHost));
  });
}

int giem2g_unet_forward_dev(void* net, const float* x_dev, float* y_dev, int B, int nx, int ny, int nz,
                            int keep_for_backward) {
  return guard([&] {
    UNet* n = NET(net);
    unet_reserve(*n, B, nx, ny, nz, keep_for_backward != 0);
    unet_load_input(*n, x_dev, n->stream);
    unet_forward(*n, n->stream);
    if (y_dev)
      G2G_HIP(hipMemcpyAsync(y_dev, n->out, sizeof(float) * 6 * (size_t)B * nx * ny * nz,
                             hipMemcpyDeviceToDevice, n->stream));
    G2G_HIP(hipStreamSynchronize(n->stream));
  });
}

int giem2g_unet_backward_dev(void* net, const float* dy_dev) {
  return guard([&] {
    UNet* n = NET(net);
    G2G_CHECK(n->training, -4, "no forward pass kept for backward");
    const size_t nv = (size_t)n->B * n->nx * n->ny * n->nz;
    // dout rows are 8 wide (channels 6, 7 zero)
    G2G_HIP(hipMemsetAsync(n->dout, 0, sizeof(float) * 8 * nv, n->stream));
    G2G_HIP(hipMemcpy2DAsync(n->dout, 8 * sizeof(float), dy_dev, 6 * sizeof(float), 6 * sizeof(float), nv,
                             hipMemcpyDeviceToDevice, n->stream));
    unet_backward(*n, n->stream);
    G2G_HIP(hipStreamSynchronize(n->stream));
  });
}

int giem2g_unet_train_step(void* net, const float* x_dev, const float* t_dev, int B, int nx, int ny, int nz,
                           float lr, double* loss) {
  return guard([&] {
    UNet* n = NET(net);
    unet_reserve(*n, B, nx, ny, nz, true);
    unet_load_input(*n, x_dev, n->stream);
    unet_forward(*n, n->stream);
    const double l = unet_mse(*n, t_dev, n->stream);
    unet_backward(*n, n->stream);
    unet_adam(*n, lr, n->stream);
    G2G_HIP(hipStreamSynchronize(n->stream));
    if (loss) *loss = l;
  });
}

int giem2g_unet_adam(void* net, float lr) {
  return guard([&] {
    UNet* n = NET(net);
    unet_adam(*n, lr, n->stream);
    G2G_HIP(hipStreamSynchronize(n->stream));
  });
}

int giem2g_op_apply_adjoint_dev(void* p, const void* in_dev, void* out_dev) {
  return guard([&] {
    Operator& op = H(p)->op;
    op_apply_adjoint(op, (const zd*)in_dev, (zd*)out_dev, op.stream);
  });
}

int giem2g_op_set_learned(void* p, void* net, double accept) {
  return guard([&] {
    Handle* h = H(p);
    if (!net) {
      h->net = nullptr;
      h->learned = PrecondHook();
      learned_free(h->lctx);
      return;
    }
    G2G_CHECK(accept > 0, -3, "accept threshold must be positive");
    h->net = NET(net);
    h->learned.apply = learned_hook;
    h->learned.ctx = &h->lctx;
    h->learned.accept = accept;
  });
}

int giem2g_op_learned_features(void* p, const void* r_dev, const void* e_dev, float* x_dev, float* t_dev) {
  return guard([&] {
    Handle* h = H(p);
    prepare_learned(*h);
    learned_features(h->lctx, (const zd*)r_dev, (const zd*)e_dev, x_dev, t_dev, h->op.stream);
    G2G_HIP(hipStreamSynchronize(h->op.stream));
  });
}

int giem2g_op_learned_apply(void* p, const void* v_dev, void* z_dev) {
  return guard([&] {
    Handle* h = H(p);
    prepare_learned(*h);
    learned_apply(h->lctx, (const zd*)v_dev, (zd*)z_dev, h->op.stream);
    G2G_HIP(hipStreamSynchronize(h->op.stream));
  });
}

// ------------------------------------------------- reference-compatible ---
void giem2g_calc_data_sizes(GIEM2G_ANOMALY_TYPE a, GIEM2G_DATA_TYPE* data) {
  const int rc = guard([&] {
    G2G_CHECK(data, -2, "null data");
    auto h = std::make_unique<Handle>();
    const long long Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
    data->tensor_size = (long long)sizeof(Handle);
    data->ie_kernel_buffer_length = Nx * Ny * 2 * Nz * (2 * Nz + 1);
    data->fft_buffers_length = 4 * Nx * Ny * 3 * Nz;
    data->giem2g_tensor = h.release();
  });
  if (rc) log_msg(std::string("giem2g_calc_data_sizes: ") + g_last_error);
}

void giem2g_prepare_ie_kernel(GIEM2G_ANOMALY_TYPE a, GIEM2G_DATA_TYPE data) {
  const int rc = guard([&] {
    Handle* h = H(data.giem2g_tensor);
    op_init(h->op, a.Nx, a.Ny, a.Nz, a.dx, a.dy, static_cast<const double*>(a.z));
  });
  if (rc) log_msg(std::string("giem2g_prepare_ie_kernel: ") + g_last_error);
}

void giem2g_calc_ie_kernel(void* p, GIEM2G_BACKGROUND_TYPE bkg, GIEM2G_ANOMALY_TYPE a, double omega) {
  const int rc = guard([&] {
    Operator& op = H(p)->op;
    // the anomaly given here defines the geometry of this build: the layer
    // of every cell (Lnumber) is derived from its borders at each call
    // (giem2g_interfaces.F90:158-166)
    G2G_CHECK(a.Nx == op.Nx && a.Ny == op.Ny && a.Nz == op.Nz, -3,
              "anomaly dimensions differ from the prepared operator's");
    if (a.z) op_set_geometry(op, a.dx, a.dy, static_cast<const double*>(a.z));
    set_background(op, bkg.Nl, static_cast<const double*>(bkg.csigb),
                   static_cast<const double*>(bkg.thickness), omega);
    build_tensor(op, 0);
  });
  if (rc) log_msg(std::string("giem2g_calc_ie_kernel: ") + g_last_error);
}

void giem2g_calc_fft_of_ie_kernel(void* p) {
  const int rc = guard([&] { build_tensor(H(p)->op, 1); });
  if (rc) log_msg(std::string("giem2g_calc_fft_of_ie_kernel: ") + g_last_error);
}

void giem2g_set_anomaly_conductivity(void* p, void* csiga) {
  const int rc = guard([&] {
    Handle* h = H(p);
    set_sigma_host(h->op, static_cast<const double*>(csiga));
    h->sigma_alias = static_cast<const double*>(csiga);
  });
  if (rc) log_msg(std::string("giem2g_set_anomaly_conductivity: ") + g_last_error);
}

void giem2g_apply_ie_operator(void* p, void* in, void* out) {
  const int rc = guard([&] {
    Handle* h = H(p);
    // the aliased conductivity may have changed since the last call
    if (h->sigma_alias) set_sigma_host(h->op, h->sigma_alias);
    apply_host(h->op, static_cast<const double*>(in), static_cast<double*>(out));
  });
  if (rc) log_msg(std::string("giem2g_apply_ie_operator: ") + g_last_error);
}

void giem2g_set_logger(void (*logger)(const char*)) { g_logger = logger; }

void giem2g_print_stats(void* p) {
  const int rc = guard([&] {
    Operator& op = H(p)->op;
    G2G_HIP(hipStreamSynchronize(op.stream));
    flush_prof(op);
    Counters& c = op.cnt;
    // the reference's split (PRINT_STATS, Integral_Equation_Module.F90:473-484,
    // and SolveEquation's, ie_solver_module.F90:98-113); times are those
    // measured (giem2g_op_time_apply, or every apply and Gram-Schmidt pass
    // while profiling is on)
    const double mv = (double)(c.matvecs + c.matvecs32);
    const double n = mv > 0 ? mv : 1.0;
    const double nd = c.dotprods > 0 ? (double)c.dotprods : 1.0;
    char buf[1024];
    std::snprintf(buf, sizeof buf,
                  "Time for multiplications: %.3e s | Time for dot products: %.3e s | "
                  "Average dot product: %.3e s | Average fft forward: %.3e s | "
                  "Average fft backward: %.3e s | Average contraction: %.3e s | Average mult: %.3e s | "
                  "Number of matrix-vector multiplications: %lld (%lld FP64, %lld complex64) | "
                  "Number of dot products: %lld in %lld Gram-Schmidt passes | tensor %.3e s, "
                  "tensor FFT %.3e s",
                  c.t_apply, c.t_dot, c.t_dot / nd, c.t_fwd / n, c.t_bwd / n, c.t_contr / n, c.t_apply / n,
                  (long long)mv, (long long)c.matvecs, (long long)c.matvecs32, (long long)c.dotprods,
                  (long long)c.gs_passes, c.t_tensor, c.t_tensor_fft);
    log_msg(buf);
    c = Counters();
  });
  if (rc) log_msg(std::string("giem2g_print_stats: ") + g_last_error);
}

void giem2g_delete_ie_operator(void* p) {
  const int rc = giem2g_op_destroy(p);
  if (rc) log_msg(std::string("giem2g_delete_ie_operator: ") + g_last_error);
}

// ---- receiver continuation (Continuation_Function_Module.F90) ----
int giem2g_rc_create(void* ie_op, int Nr, const double* xshift, const double* yshift, const double* depth,
                     void** out) {
  return guard([&] {
    G2G_CHECK(out != nullptr && xshift && yshift && depth, -2, "null argument");
    auto h = std::make_unique<RcHandle>();
    Handle* owner = H(ie_op);
    try {
      rc_init(h->rc, owner->op, Nr, xshift, yshift, depth);
    } catch (...) {
      rc_fini(h->rc);
      throw;
    }
    h->owner = owner;
    owner->receivers++;
    *out = h.release();
  });
}

int giem2g_rc_destroy(void* p) {
  return guard([&] {
    if (!p) return;
    RcHandle* h = static_cast<RcHandle*>(p);
    rc_fini(h->rc);
    if (h->owner) h->owner->receivers--;
    delete h;
  });
}

int giem2g_rc_calc_kernel(void* p) {
  return guard([&] { rc_build_tensor(RC(p)); });
}

int giem2g_rc_calc_fft_of_kernel(void* p) {
  return guard([&] { rc_tensor_fft(RC(p)); });
}

int giem2g_rc_get_tensor(void* p, double* G) {
  return guard([&] {
    RcOperator& rc = RC(p);
    G2G_CHECK(rc.have_tensor, -4, "receiver tensor not built");
    G2G_HIP(hipMemcpy(G, rc.G, sizeof(zd) * rc.NP * rc.BS, hipMemcpyDeviceToHost));
  });
}

int giem2g_rc_put_tensor(void* p, const double* G, int spectral) {
  return guard([&] {
    RcOperator& rc = RC(p);
    G2G_HIP(hipMemcpy(rc.G, G, sizeof(zd) * rc.NP * rc.BS, hipMemcpyHostToDevice));
    rc_prepare_receivers(rc);
    rc.have_tensor = true;
    rc.spectral = spectral != 0;
    rc.tensor_version = rc.op->bkg_version;
  });
}

int giem2g_rc_receivers(void* p, double* zr, int* layer, int* cell) {
  return guard([&] {
    RcOperator& rc = RC(p);
    rc_prepare_receivers(rc);
    for (int r = 0; r < rc.Nr; r++) {
      if (zr) zr[r] = rc.zr[r];
      if (layer) layer[r] = rc.layer[r];
      if (cell) cell[r] = rc.cell[r];
    }
  });
}

int giem2g_rc_set_anomaly_conductivity(void* p, const double* csiga) {
  return guard([&] {
    RcOperator& rc = RC(p);
    DevCopy d(csiga, (size_t)rc.op->ncell());
    rc_set_sigma(rc, d.d);
  });
}

int giem2g_rc_set_anomaly_conductivity_dev(void* p, const void* csiga) {
  return guard([&] { rc_set_sigma(RC(p), static_cast<const zd*>(csiga)); });
}

int giem2g_rc_recalculate_dev(void* p, const void* Eint, void* Ea, void* Ha) {
  return guard([&] {
    RcOperator& rc = RC(p);
    rc_apply(rc, static_cast<const zd*>(Eint), static_cast<zd*>(Ea), static_cast<zd*>(Ha), rc.op->stream);
  });
}

int giem2g_rc_recalculate(void* p, const double* Eint, double* Ea, double* Ha) {
  return guard([&] {
    RcOperator& rc = RC(p);
    const Operator& op = *rc.op;
    const size_t no = (size_t)rc.Nr * 3 * op.Nx * op.Ny;
    DevCopy e(Eint, (size_t)op.nvec()), a(nullptr, no), h(nullptr, no);
    rc_apply(rc, e.d, a.d, h.d, op.stream);
    G2G_HIP(hipMemcpy(Ea, a.d, sizeof(zd) * no, hipMemcpyDeviceToHost));
    G2G_HIP(hipMemcpy(Ha, h.d, sizeof(zd) * no, hipMemcpyDeviceToHost));
  });
}

int giem2g_rc_plane_wave(void* p, int pol, double* F0) {
  return guard([&] {
    RcOperator& rc = RC(p);
    rc_prepare_receivers(rc);
    plane_wave_at(*rc.op, pol, rc.Nr, rc.zr.data(), F0);
  });
}

int giem2g_rc_stats(void* p, double* o) {
  return guard([&] {
    RcOperator& rc = RC(p);
    Operator& op = *rc.op;
    float ms[4] = {0, 0, 0, 0};
    if (rc.applies > 0) {
      G2G_HIP(hipEventSynchronize(op.ev[6]));
      G2G_HIP(hipEventElapsedTime(&ms[0], op.ev[4], op.ev[6]));
      G2G_HIP(hipEventElapsedTime(&ms[1], op.ev[4], op.ev[7]));
      G2G_HIP(hipEventElapsedTime(&ms[2], op.ev[7], op.ev[5]));
      G2G_HIP(hipEventElapsedTime(&ms[3], op.ev[5], op.ev[6]));
    }
    const double v[8] = {rc.t_tensor, rc.t_fft, ms[0], ms[1], ms[2], ms[3], (double)rc.applies,
                         (double)rc.bytes};
    std::memcpy(o, v, sizeof v);
  });
}

}  // extern "C"
