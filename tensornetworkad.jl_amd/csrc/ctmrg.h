// Host-side orchestration types of the CTMRG step, its reverse sweep and the
// energy chain (ctmrg.hip).
#pragma once
#include <vector>

#include "tc_build.h"
#include "tnad_internal.h"

namespace tnad {

// buffer at a contraction precision (double for F64, float for F32 / BF16)
struct PBuf {
  Prec p = Prec::F64;
  size_t n = 0;
  DevBuf b;
  PBuf() = default;
  PBuf(Ctx* c, Prec p_, size_t n_);
  PBuf(PBuf&&) = default;
  PBuf& operator=(PBuf&&) = default;
  void* ptr() const { return b.ptr; }
};
void to_prec(Ctx* c, const double* src, PBuf& dst);
void from_prec(Ctx* c, const PBuf& src, double* dst);

// one contraction launch from einsum-style views
void tc(Ctx* c, Prec p, const TView& A, const void* a, const TView& B, const void* b, const TView& C, void* o,
        double alpha = 1.0, double beta = 0.0);

struct StepDims {
  int chi = 0, D = 0, n = 0;
};

struct StepOpts {
  Prec prec = Prec::F64;      // contraction precision of this step
  int eig_block = 0;          // subspace block size k (n > 192); 0 = automatic
  int eig_max_iters = 200;    // cap, cold start
  int eig_warm_iters = 200;   // cap, warm start (previous step's Ritz vectors)
  double eig_tol = 1e-13;     // relative Ritz residual of the kept vectors (FP64 steps)
  double eig_tol_low = 1e-13; // the same for steps at a lower contraction precision
};

struct StepState {
  int k = 0;
  int eig_iters = 0;
  double eig_resid = 0.0;
  // n > 192 only: norm of the singular values beyond the chi kept ones, relative to s[1]
  // (|M - Z Theta Z^T|_F / s[1]), and a bound on its rounding and deflation error
  double rest = 0.0, rest_err = 0.0;
  double eig_ms = 0.0;   // device time of this step's eigensolver
};

struct EigStats {
  int iters = 0, k = 0;
  double resid = 0.0;
  int exact = 0;    // Rayleigh-Ritz steps solved exactly (the selective path declined)
};
struct EigBackStats {
  int pcg_iters = 0;
};

struct EigWork {
  DevBuf Y, Xr, Yr, H, W, Vw, G, Rinv, nrm, res, info, scal;
  EigWork(Ctx* c, int n, int k);
};

// Top-k eigenpairs (by |lambda|) of the symmetric n x n FP64 matrix A.
// n <= 192: full Jacobi (k = n).  Otherwise block subspace iteration with
// Rayleigh-Ritz and CholeskyQR orthonormalisation (one pass, CholeskyQR2 or shifted CholeskyQR3 as the
// block needs), optionally warm-started from X0.
// Outputs X (n x k), AX = A X, theta (k).  Returns k.
int topk_eig(Ctx* c, const double* A, int n, int want, int k, const double* X0, int maxit, double tol,
             double* X, double* AX, double* theta, EigStats* st);

// Subspace block of a step: n (full spectrum) for n <= 192; else the chi kept columns plus a guard of
// 1.25 chi (2.25 chi in all), or eig_block when set.  The guard sets the convergence rate
// |lambda_(k+1) / lambda_chi| per iteration: on random normal D=3 bulks at chi=64 a block of 96
// (1.5 chi) took ~40 iterations per step, 144 about 2.5x fewer and 25% less time.  Capped at 192 (the
// exact Rayleigh-Ritz solver's limit); at chi = 128, D = 4 the energy + gradient runs 3% (mixed
// precision) and 4% (FP64) faster at k = 192 than at 176, and 6% / 18% faster than at 160.
inline int step_k(const StepDims& d, const StepOpts& o) {
  if (d.n <= JACOBI_MAX) return d.n;
  int k = o.eig_block > 0 ? o.eig_block : 2 * d.chi + d.chi / 4;
  if (k < d.chi) k = d.chi;
  if (k > JACOBI_MAX) k = JACOBI_MAX;
  if (k > d.n) k = d.n;
  return k;
}

// checkpoint of one step for the reverse sweep
struct StepTape {
  Prec prec;
  int k;
  double eig_tol = 1e-13;   // relative Ritz residual the step's eigensolver was run to
  DevBuf C, T;          // step inputs (FP64)
  DevBuf X, AX, theta;  // eigenpairs (FP64): X = [z | extra], AX = A X
  DevBuf M;             // cpmat + cpmat^T (FP64)
  PBuf X2, Y2;          // contraction intermediates reused by the adjoint
  DevBuf C2, T2, nrm;   // symmetrised, unnormalised outputs and their norms
  StepTape(Ctx* c, const StepDims& d, Prec p, int k);
};

// X0: warm-start block (n x k) or null.  X_keep: if non-null, receives this step's
// Ritz vectors (n x k) for the next step's warm start.
void ctmrg_step(Ctx* c, const StepDims& d, const StepOpts& o, const double* a64, const PBuf& ap,
                const double* C, const double* T, const double* X0, double* X_keep, double* Cout, double* Tout,
                std::vector<double>& vals, StepTape* tape, StepState& state);

// sweep_ref (nullable): running reference of the reverse sweep for the resolvent solve (see
// resolvent_pcg); a step computed at a lower contraction precision is solved to the same absolute
// accuracy as the sweep's largest right-hand side.
void eig_adjoint(Ctx* c, const StepDims& d, const StepTape& tp, const double* G, const double* C1bar,
                 double* cpbar, EigBackStats* st, double* sweep_ref = nullptr);

void ctmrg_step_back(Ctx* c, const StepDims& d, Prec p, const StepTape& tp, const double* a64, const PBuf& ap,
                     const double* C3bar, const double* T3bar, double* Cbar, double* Tbar, double* abar,
                     double* sweep_ref = nullptr);

struct EvCache {
  DevBuf apn, E1, E2, E3, E4, l, tr, scal;
  EvCache(Ctx* c, int chi, int D, int ph);
};
void expectation_fwd(Ctx* c, int chi, int D, int ph, const double* h, const double* ap, const double* C,
                     const double* T, EvCache& ev, double* e_host);
void expectation_back(Ctx* c, int chi, int D, int ph, const double* h, const double* ap, const double* C,
                      const double* T, EvCache& ev, double* apbar, double* Cbar, double* Tbar);

}  // namespace tnad
