// g2g_solver.hpp -- nested FGMRES interface (see g2g_solver.hip).
#pragma once
#include <functional>
#include <vector>

#include "g2g_op.hpp"

namespace g2g {

// Arena slots of one solve's device buffers (Operator::ws): the outer and
// inner solvers' vectors and bases, the mixed-precision and learned-step
// vectors, and the BLAS reduction buffers
enum WsSlot {
  WS_OUTER = 0,   // 5 slots: x, r0, w, V, Z
  WS_IN64 = 5,    // 4 slots: x, r0, w, V
  WS_IN32 = 9,    // 4 slots
  WS_B32 = 13, WS_X32, WS_R32, WS_TMP, WS_TMP2,
  WS_PART, WS_RES, WS_HBUF,
  WS_RHS, WS_GUESS, WS_PAIR  // the C ABI's solve entry points
};


struct SolveParams {
  int outer_m = 30, inner_m = 30, max_restart = 100;
  double tol = 1e-8;
  int inner_precision = 0;  // 0 FP64, 1 complex64
  int precond = 0;          // 0 inner GMRES, 1 learned + inner GMRES fallback
  // learned modes: after this many rejected learned steps in a row the
  // network is not called again in the solve (0: never stop)
  int learned_patience = 2;
  // inner solves stop below this residual relative to their right-hand
  // side; < 0: relaxed, |inner_tol| tol divided by the outer's current
  // relative residual, between tol and 0.1; 0: tol for FP64 inner solves
  // (the reference's inner and outer share one tolerance), -1 for complex64
  double inner_tol = 0;
};

struct SolveResult {
  int iterations = 0;
  int stat = -1;
  double bea = -1, be = -1;
  long long precond_calls = 0, precond_fallbacks = 0;
  // Krylov work of the solve: dot products, Gram-Schmidt passes, bytes moved
  long long dots = 0, passes = 0;
  double krylov_bytes = 0;
};

// A learned right preconditioner: z = M(v) on HBM vectors (complex128).
struct PrecondHook {
  void (*apply)(void* ctx, const zd* v, zd* z) = nullptr;
  void* ctx = nullptr;
  double accept = 1.0;  // accept the learned step when ||v - A M(v)|| < accept*||v||
};

// Where a solve runs: the stream of its Krylov BLAS, its matrix-vector
// products and their counters (default: the operator's stream and apply).
struct SolveEnv {
  hipStream_t stream = nullptr;
  std::function<void(const zd*, zd*)> A64;
  std::function<void(const zf*, zf*)> A32;
  int64_t* mv64 = nullptr;
  int64_t* mv32 = nullptr;
  Arena* ws = nullptr;  // the solve's device buffers
};

SolveResult nested_solve(Operator& op, const zd* b, const zd* x0, zd* xout,
                         const SolveParams& p, std::vector<double>* hist, const PrecondHook* learned,
                         const SolveEnv* env = nullptr);

struct PairResult {
  SolveResult r[2];
  int64_t mv64[2] = {0, 0}, mv32[2] = {0, 0};
  std::vector<double> hist[2];
  long long paired = 0, alone = 0;  // two-vector and one-vector operator applications
};
// two independent solves (b[0], b[1]) whose operator products are paired
PairResult nested_solve_pair(Operator& op, const zd* const b[2], const zd* const x0[2], zd* const xout[2],
                             const SolveParams& p);
// mode 0: out = sqrt(Re sb)*in (rhs); 1: out = in/asiga (warm start); 2: out = asiga*in (field)
void solve_scale(Operator& op, zd* out, const zd* in, int mode);

}  // namespace g2g
