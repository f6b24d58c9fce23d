// Shifted resolvent solves of the truncated eigen / singular-value adjoints (resolvent.hip).
#pragma once
#include <functional>

#include "tnad_internal.h"

namespace tnad {

// Solves, for every kept column j < r,
//     (sig_j - G) x_j = P b_j      on the range of P = I - Wh Wh^T,
// where Wh = W[:, 0:r] are the kept (top) eigenvectors of the symmetric positive
// semi-definite operator G and sig_j their eigenvalues, so the system is positive
// definite on that range.  W[:, r:k] (with estimates tz of their G-eigenvalues) are
// the further vectors the eigensolver computed; they need not be converged: they only
// precondition the conjugate-gradient iteration
//     M_j^-1 = Wt L(sig_j - tz) Wt^T + (I - W W^T) / sig_j,   L(d) = d / (d^2 + eta),
// which is exact when they are eigenvectors (then one step suffices) and otherwise
// leaves a spectrum in [1 - g_max / sig_j, 1] on the rest.  L is the reference's
// Lorentzian broadening of 1/(s_j^2 - s_i^2) (src/trg.jl:72-78), so an exactly
// degenerate pair across the cut contributes nothing, as there.
// The iteration stops per column when the preconditioned residual r.M^-1 r falls
// below tol^2 times b.M^-1 b, or below 1e-30 |b_j|^2 / sig_j (the complement part of b
// is then at the rounding level of the whole).  applyG(V, GV) must write G V for a
// rows x r block V.
// sweep_ref (nullable, host): the largest b.M^-1 b of any column in the solves so far (a reverse
// sweep's running reference); this solve raises it to its own largest.  With abs_tol > 0 a column
// also stops once r.M^-1 r <= abs_tol^2 * sweep_ref: the same absolute accuracy as the sweep's
// largest right-hand side, for a solve whose right-hand side is much smaller.
// All pointers are device FP64 arrays, column-major; sig (r) and tz (k - r) on device.
// Returns the number of iterations.
int resolvent_pcg(Ctx* c, int rows, int r, int k, const double* W, const double* sig, const double* tz,
                  const std::function<void(const double*, double*)>& applyG, const double* b, double* x,
                  double tol, int maxit, double eta = 1e-40, double* sweep_ref = nullptr, double abs_tol = 0.0);

// out[i] = in[i]^2
void k_square(const double* in, double* out, int n, hipStream_t s);

}  // namespace tnad
