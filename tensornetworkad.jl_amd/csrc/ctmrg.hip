// CTMRG renormalisation step, its reverse sweep, the top-k FP64 eigensolver
// and the energy chain, orchestrated on one HIP stream.
//
// Reference: src/ctmrg.jl:110-153 (ctmrg, ctmrgstep), src/fixedpoint.jl:11-41
// (fixedpoint, StopFunction), src/trg.jl:72-106 (svd_back), src/autodiff.jl:23-29
// (norm adjoint), src/variationalipeps.jl:28-56 (energy, expectationvalue).
//
// Notation: chi = environment bond, D = double-layer bond (Dd = d^2 in the
// reference), n = chi*D.  Every tensor is column-major with the reference's
// index order; contractions are issued as the reference's einsum strings
// through contract_desc (tc_build.cpp).
//
// The n x n "corner matrix" cpmat + cpmat^T is symmetric, so its SVD is
// obtained from an eigendecomposition: s = |lambda| sorted descending,
// U = eigenvectors, V = U sign(lambda).
#include "ctmrg.h"
#include "resolvent.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace tnad {

static inline size_t esize(Prec p) { return p == Prec::F64 ? 8 : 4; }

// --------------------------------------------------------------------------
// small kernels private to this file
// --------------------------------------------------------------------------
namespace {

constexpr int NT = 256;

// out[j] = |Y[:, j] - theta_j X[:, j]| for j < gridDim.x; block 0 also writes out[gridDim.x] = theta_0 and
// out[gridDim.x + 1] = the Rayleigh-Ritz flag (*gate, or 1 without one), so that the host reads the
// iteration's whole status with one transfer
__global__ void resid_norms_kernel(const double* __restrict__ X, const double* __restrict__ Y,
                                   const double* __restrict__ theta, int n, double* __restrict__ out,
                                   const int* __restrict__ gate) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  if (j == 0 && threadIdx.x == 0) {
    out[gridDim.x] = theta[0];
    out[gridDim.x + 1] = gate ? (double)*gate : 1.0;
  }
  const double th = theta[j];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double r = Y[i + (long long)j * n] - th * X[i + (long long)j * n];
    acc += r * r;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// Next subspace block from the Rayleigh-Ritz pairs, and its column norms, in one pass:
// Z[:, j] = Yr[:, j] / theta_j = Xr[:, j] + R[:, j] / theta_j (R = Yr - Xr Theta is orthogonal to Xr, so
// Z^T Z = I + (R Theta^-1)^T (R Theta^-1) >= I stays well conditioned however collinear the residuals are);
// numerically null Ritz values keep Z[:, j] = Xr[:, j].  One block of 256 threads per column j, the sum of
// squares taken in the order of the column-norm kernel (thread-strided partial sums, a shuffle tree per
// wave, the wave sums added in order).
__global__ __launch_bounds__(256) void ritz_scale_norms_kernel(const double* __restrict__ Xr,
                                                               const double* __restrict__ Yr,
                                                               const double* __restrict__ theta, int n,
                                                               double* __restrict__ Z, double* __restrict__ nrm) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  const double t0 = fabs(theta[0]), th = theta[j];
  const bool use_y = fabs(th) > 1e-14 * t0;
  const size_t o = (size_t)n * j;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double z = use_y ? Yr[o + i] / th : Xr[o + i];
    Z[o + i] = z;
    acc += z * z;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int w = 0; w < 4; ++w) sum += red[w];
    nrm[j] = sqrt(sum);
  }
}

// *out = sum_ij (G_ij - delta_ij)^2 (one workgroup; G is k x k)
// |G - I|_F^2 of a k x k matrix (k <= 256): one workgroup of GD_NT threads, thread (row i, column
// phase jo) sums rows i over columns jo, jo + GD_NT / 256, ... (coalesced along rows)
constexpr int GD_NT = 1024;
// out[0] = |G - I|_F^2; out[1] = *info (the preceding Cholesky's breakdown column, as a double), so the host
// reads the pass's status with the pivot in one transfer
__global__ __launch_bounds__(GD_NT) void gram_dev_kernel(const double* __restrict__ G, int k,
                                                         double* __restrict__ out, const int* __restrict__ info) {
  __shared__ double red[GD_NT / 64];
  const int t = threadIdx.x, i = t & 255, jo = t >> 8;
  double acc = 0.0;
  if (i < k)
    for (int j = jo; j < k; j += GD_NT / 256) {
      const double v = G[i + (size_t)k * j] - (i == j ? 1.0 : 0.0);
      acc += v * v;
    }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((t & 63) == 0) red[t >> 6] = acc;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int w = 0; w < GD_NT / 64; ++w) s += red[w];
    out[0] = s;
    out[1] = (double)*info;
  }
}

// F = (I + E)^{-1/2} to second order for G = I + E (binomial series: I - E/2 + 3/8 E^2), written from G and
// G2 = G G as 15/8 I - 5/4 G + 3/8 G2; F may overwrite G2 (each entry is read before it is written)
__global__ void inv_sqrt2_coef_kernel(const double* __restrict__ G, const double* G2, int k, double* F) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k * k) return;
  const int i = q % k, j = q / k;
  F[q] = 0.375 * G2[q] - 1.25 * G[q] + (i == j ? 1.875 : 0.0);
}

// make the largest-|.| entry of every column positive (first occurrence wins);
// the same sign is applied to the companion block Y (A*X) when given.
__global__ void canon_sign_kernel(double* __restrict__ X, double* __restrict__ Y, int n) {
  __shared__ double bv[4];
  __shared__ int bi[4];
  const int j = blockIdx.x;
  double best = -1.0;
  int idx = 0x7fffffff;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double v = fabs(X[i + (long long)j * n]);
    if (v > best || (v == best && i < idx)) { best = v; idx = i; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    double ob = __shfl_down(best, o);
    int oi = __shfl_down(idx, o);
    if (ob > best || (ob == best && oi < idx)) { best = ob; idx = oi; }
  }
  if ((threadIdx.x & 63) == 0) { bv[threadIdx.x >> 6] = best; bi[threadIdx.x >> 6] = idx; }
  __syncthreads();
  __shared__ double sgn;
  if (threadIdx.x == 0) {
    double b = bv[0];
    int ii = bi[0];
    for (int w = 1; w < 4; ++w)
      if (bv[w] > b || (bv[w] == b && bi[w] < ii)) { b = bv[w]; ii = bi[w]; }
    sgn = X[ii + (long long)j * n] < 0 ? -1.0 : 1.0;
  }
  __syncthreads();
  if (sgn < 0) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      X[i + (long long)j * n] = -X[i + (long long)j * n];
      if (Y) Y[i + (long long)j * n] = -Y[i + (long long)j * n];
    }
  }
}

// coefficient block of the eigen adjoint among the kept pairs (see eig_adjoint below):
//   coef = Kt/2 with Kt = R11 Sigma + Sigma R11^T + C1bar,  R11_ij = (F_ij g_ij + F_ji g_ji) s_j,
// F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eta),  s = |lambda|.
// Evaluated without FP contraction: the reference forms S2 = S.^2 first and then
// S2' .- S2 (src/trg.jl:76-78), so the diagonal difference is exactly zero and the
// Lorentzian F vanishes there; a fused multiply-add would leave the rounding residual of
// s^2 (~1e-17) in place of that zero and F would blow up to ~1e23.
__global__ void eigback_coef_kernel(const double* __restrict__ g, const double* __restrict__ lam,
                                    const double* __restrict__ C1bar, int chi, double eta,
                                    double* __restrict__ coef) {
#pragma clang fp contract(off)
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= chi * chi) return;
  int i = e % chi, j = e / chi;
  double li = lam[i], lj = lam[j];
  double si2 = li * li, sj2 = lj * lj;
  double Fij = (sj2 - si2) / ((sj2 - si2) * (sj2 - si2) + eta);
  double Fji = -Fij;
  double gij = g[i + (long long)j * chi];
  double gji = g[j + (long long)i * chi];
  double sj = fabs(lj), si = fabs(li);
  double sigj = lj < 0 ? -1.0 : 1.0, sigi = li < 0 ? -1.0 : 1.0;
  double R_ij = (Fij * gij + Fji * gji) * sj;   // R11[i,j]
  double R_ji = (Fji * gji + Fij * gij) * si;   // R11[j,i]
  double Kt = R_ij * sigj + sigi * R_ji + C1bar[i + (long long)j * chi];
  coef[e] = 0.5 * Kt;
}

// W[:, j] += lambda_j * Xr[:, j] + AXr[:, j]
__global__ void add_lam_cols_kernel(double* __restrict__ W, const double* __restrict__ Xr,
                                    const double* __restrict__ AXr, const double* __restrict__ lam, int n,
                                    int chi) {
  size_t tot = (size_t)n * chi;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    int j = e / n;
    W[e] += lam[j] * Xr[e] + AXr[e];
  }
}

// lbar[abc, x, y] += coef * tr[abc] * delta_xy
__global__ void add_trace_diag_kernel(double* __restrict__ lbar, const double* __restrict__ tr,
                                      const double* __restrict__ coef_dev, double coef_host, int m, int s) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * s) return;
  int abc = e % m, x = e / m;
  double c = coef_dev ? *coef_dev : coef_host;
  lbar[abc + (size_t)m * (x + s * x)] += c * tr[abc];
}

// tr[abc] = sum_x l[abc, x, x]
__global__ void trace_last2_kernel(const double* __restrict__ l, int m, int s, double* __restrict__ tr) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  double acc = 0.0;
  for (int x = 0; x < s; ++x) acc += l[e + (size_t)m * (x + s * x)];
  tr[e] = acc;
}

// out = a / b  and the energy-adjoint scalars (ebar = 1/nn, nnbar = -e/nn^2), all on device
__global__ void energy_scalars_kernel(const double* e, const double* nn, double* out) {
  out[0] = e[0] / nn[0];
  out[1] = 1.0 / nn[0];
  out[2] = -e[0] / (nn[0] * nn[0]);
}

// norm adjoint: out = ybar/|x| + nbar * x / (|x| + eps(0f0)),  nbar = -<x,ybar>/|x|^2
__global__ void norm_back_kernel(const double* __restrict__ x, const double* __restrict__ ybar,
                                 const double* __restrict__ sumsq, const double* __restrict__ dot,
                                 double* __restrict__ out, size_t n) {
  const double nrm = sqrt(*sumsq);
  const double nbar = -(*dot) / (nrm * nrm);
  const double eps0 = 1.401298464324817e-45;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = ybar[i] / nrm + nbar * x[i] / (nrm + eps0);
}

__global__ void sqrt_inplace_kernel(double* x) { x[0] = sqrt(x[0]); }

inline unsigned nb(size_t n) {
  size_t b = (n + NT - 1) / NT;
  if (b > 65535ull * 8) b = 65535ull * 8;
  return (unsigned)(b ? b : 1);
}

}  // namespace

// --------------------------------------------------------------------------
// typed buffer helpers
// --------------------------------------------------------------------------
PBuf::PBuf(Ctx* c, Prec p_, size_t n_) : p(p_), n(n_), b(c, n_ * esize(p_)) {}

void to_prec(Ctx* c, const double* src, PBuf& dst) {
  if (dst.p == Prec::F64) k_copy(src, dst.b.as<double>(), dst.n, c->stream);
  else k_cast_from<float>(src, dst.b.as<float>(), dst.n, c->stream);
}
void from_prec(Ctx* c, const PBuf& src, double* dst) {
  if (src.p == Prec::F64) k_copy(src.b.as<double>(), dst, src.n, c->stream);
  else k_cast<float>(src.b.as<float>(), dst, src.n, c->stream);
}

// src at precision p: the buffer itself when the storage type already matches,
// otherwise a converted copy held in tmp (FP64 <-> float storage)
const PBuf& at_prec(Ctx* c, const PBuf& src, Prec p, PBuf& tmp) {
  const bool src64 = src.p == Prec::F64, want64 = p == Prec::F64;
  if (src64 == want64) return src;
  tmp = PBuf(c, p, src.n);
  if (src64) {
    to_prec(c, src.b.as<double>(), tmp);
  } else {
    k_cast<float>(src.b.as<float>(), tmp.b.as<double>(), src.n, c->stream);
  }
  return tmp;
}

void tc(Ctx* c, Prec p, const TView& A, const void* a, const TView& B, const void* b, const TView& C, void* o,
        double alpha, double beta) {
  tc_launch(c, p, contract_desc(A, B, C, alpha, beta), a, b, o, c->stream);
}

// Opt-in diagnostics (TNAD_DEBUG_EIG=1): report whether a device array holds a non-finite value.
static void dbg_finite(Ctx* c, const char* what, const double* dev, size_t n) {
  static const bool on = [] {
    const char* e = std::getenv("TNAD_DEBUG_FINITE");
    return e && *e && *e != '0';
  }();
  if (!on) return;
  std::vector<double> h(n);
  c->read_async(h.data(), dev, n * 8);
  c->sync();
  size_t bad = 0;
  double mx = 0.0;
  for (double v : h) {
    if (!std::isfinite(v)) ++bad;
    else mx = std::max(mx, std::fabs(v));
  }
  std::fprintf(stderr, "[tnad eig]   %s: %zu non-finite of %zu, max |finite| %.3e\n", what, bad, n, mx);
}

// ==========================================================================
// Eigensolver: top-k eigenpairs of a symmetric n x n FP64 matrix by |lambda|
// ==========================================================================
// norms_ready: w.nrm already holds the column norms of Y (ritz_scale_norms_kernel)
static void orthonormalize(Ctx* c, double* Y, int n, int k, double* X, EigWork& w, unsigned seed,
                           bool norms_ready = false) {
  hipStream_t s = c->stream;
  // column scaling, random refill of (numerically) null columns
  if (norms_ready) k_fill_scale_cols(Y, n, k, w.nrm.as<double>(), 1e-13, seed, s);
  else k_normalize_cols(Y, n, k, w.nrm.as<double>(), 1e-13, seed, s);
  // A well-conditioned block takes plain CholeskyQR passes: one when the result is orthonormal to
  // working precision (|X^T X - I|_F <= 1e-12, the usual case once the iteration converges), else a
  // second: X (X^T X)^{-1/2} to second order when |X^T X - I|_F <= 1e-5, CholeskyQR2 above.  That holds when the Cholesky pivots of the Gram matrix stay above 1e-4 (unit
  // columns: condition number around 1e4, far below u^-1/2).  The first pass and the test run before
  // the host looks at the pivots, so that one round trip serves both.  Otherwise: shifted CholeskyQR
  // followed by two plain passes (sCholQR3); columns that the shifted pass finds linearly dependent
  // on earlier ones (pivot at the shift's level: the block has lower rank than k, e.g. a
  // rank-deficient corner matrix) are replaced by random directions first, so the block always spans
  // k dimensions.
  std::vector<double> piv(k);
  std::vector<int> dead(k);
  const double u = 1.1102230246251565e-16;
  const double shift = 11.0 * ((double)n * k + (double)k * (k + 1)) * u * (double)k;
  // a column with less than 1% of its (unit) norm outside the span of the earlier ones is replaced:
  // it adds next to no new direction, and keeping it would leave the block so ill-conditioned that
  // plain passes could break down
  const double dead_thresh = std::max(2.0 * std::sqrt(shift), 1e-2);
  // shifted Cholesky of the Gram matrix of Y: its pivots flag the dependent columns (returns how many)
  auto find_dead = [&](unsigned fill_seed, bool refill) {
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', k, k, n, n, n, k), Y, Y, w.G.as<double>(), s);
    chol_inv(w.G.as<double>(), k, shift, w.Rinv.as<double>(), w.info.as<int>(), w.res.as<double>(), s);
    c->read_async(piv.data(), w.res.ptr, (size_t)k * 8);
    c->sync();
    int ndead = 0;
    for (int j = 0; j < k; ++j) {
      dead[j] = !(piv[j] > dead_thresh);
      ndead += dead[j];
    }
    if (ndead > 0 && refill) {
      c->write_async(w.info.ptr, dead.data(), (size_t)k * sizeof(int));
      k_random_fill_flagged(Y, n, k, w.info.as<int>(), fill_seed, s);
      k_col_norms(Y, n, k, w.nrm.as<double>(), s);
      k_scale_cols(Y, n, k, w.nrm.as<double>(), 1, s);
    }
    return ndead;
  };
  // Plain CholeskyQR as described above; when its pivots are too small, the shifted Cholesky finds
  // the dependent columns, they are refilled, and the plain passes are tried once more (a block whose
  // only defect was those columns is now well-conditioned).  What still fails takes sCholQR3 below.
  for (int round = 0; round < 2; ++round) {
    double* scal = w.scal.as<double>();   // [0] smallest pivot, [1] |X^T X - I|_F^2, [2] breakdown column
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', k, k, n, n, n, k), Y, Y, w.G.as<double>(), s);
    chol_inv(w.G.as<double>(), k, 0.0, w.Rinv.as<double>(), w.info.as<int>(), nullptr, s, scal);
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, k, n, k, n), Y, w.Rinv.as<double>(), X, s);
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', k, k, n, n, n, k), X, X, w.G.as<double>(), s);
    hipLaunchKernelGGL(gram_dev_kernel, dim3(1), dim3(GD_NT), 0, s, w.G.as<double>(), k, scal + 1,
                       w.info.as<int>());
    TNAD_HIP_TRY(hipGetLastError());
    double hs[3] = {0.0, 0.0, 0.0};   // smallest pivot, |X^T X - I|_F^2, Cholesky breakdown column
    c->read_async(hs, scal, 3 * sizeof(double));
    c->sync();
    int info = (int)hs[2];
    if (std::getenv("TNAD_DEBUG_EIG"))
      std::fprintf(stderr, "[tnad eig] plain Cholesky info=%d min pivot=%.3e |X^T X - I|_F=%.3e\n", info, hs[0],
                   std::sqrt(hs[1]));
    if (info == 0 && hs[0] > 1e-4) {
      const double dev = std::sqrt(hs[1]);
      if (dev <= 1e-12) return;
      if (dev <= 1e-5) {
        // X^T X = I + E with |E|_F <= 1e-5: X F with F = (I + E)^{-1/2} to second order is orthonormal to
        // O(|E|^3) (below 1e-15) and spans the same columns, which is all the iteration keeps; it takes
        // two small products instead of a second Cholesky and its host check
        tc_launch(c, Prec::F64, gemm_desc('N', 'N', k, k, k, k, k, k), w.G.as<double>(), w.G.as<double>(),
                  w.Rinv.as<double>(), s);
        hipLaunchKernelGGL(inv_sqrt2_coef_kernel, dim3((k * k + 255) / 256), dim3(256), 0, s, w.G.as<double>(),
                           w.Rinv.as<double>(), k, w.Rinv.as<double>());
        TNAD_HIP_TRY(hipGetLastError());
        tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, k, n, k, n), X, w.Rinv.as<double>(), Y, s);
        k_copy(Y, X, (size_t)n * k, s);
        return;
      }
      chol_inv(w.G.as<double>(), k, 0.0, w.Rinv.as<double>(), w.info.as<int>(), nullptr, s);
      c->read_async(&info, w.info.ptr, sizeof(int));
      c->sync();
      if (info != 0)   // breakdown (pivot <= 0): factor the shifted Gram instead
        chol_inv(w.G.as<double>(), k, shift, w.Rinv.as<double>(), w.info.as<int>(), nullptr, s);
      tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, k, n, k, n), X, w.Rinv.as<double>(), Y, s);
      k_copy(Y, X, (size_t)n * k, s);
      return;
    }
    if (round == 1) break;
    const int ndead = find_dead(seed + 7919u, true);
    if (std::getenv("TNAD_DEBUG_EIG"))
      std::fprintf(stderr, "[tnad eig] shifted CholeskyQR: %d dependent columns refilled\n", ndead);
    if (ndead == 0) break;   // ill-conditioned without dependent columns: sCholQR3
  }
  // the Cholesky factor used below must belong to the block as it finally is: the last attempt
  // factors without refilling afterwards
  for (int attempt = 0;; ++attempt) {
    const int ndead = find_dead(seed + 7919u * (attempt + 2), attempt < 3);
    if (std::getenv("TNAD_DEBUG_EIG"))
      std::fprintf(stderr, "[tnad eig] shifted CholeskyQR attempt %d: %d dependent columns\n", attempt, ndead);
    if (ndead == 0 || attempt == 3) break;
  }
  dbg_finite(c, "shifted Rinv", w.Rinv.as<double>(), (size_t)k * k);
  tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, k, n, k, n), Y, w.Rinv.as<double>(), X, s);
  double* src = X;
  double* dst = Y;
  for (int pass = 1; pass < 3; ++pass) {
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', k, k, n, n, n, k), src, src, w.G.as<double>(), s);
    chol_inv(w.G.as<double>(), k, 0.0, w.Rinv.as<double>(), w.info.as<int>(), nullptr, s);
    int info = 0;
    c->read_async(&info, w.info.ptr, sizeof(int));
    c->sync();
    if (info != 0)   // breakdown (pivot <= 0): factor the shifted Gram instead
      chol_inv(w.G.as<double>(), k, shift, w.Rinv.as<double>(), w.info.as<int>(), nullptr, s);
    dbg_finite(c, "plain pass Rinv", w.Rinv.as<double>(), (size_t)k * k);
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, k, n, k, n), src, w.Rinv.as<double>(), dst, s);
    std::swap(src, dst);
  }
  // passes alternate X -> Y -> X: the orthonormal block ends in X (Y is scratch)
}

EigWork::EigWork(Ctx* c, int n, int k)
    : Y(c, (size_t)n * k * 8), Xr(c, (size_t)n * k * 8), Yr(c, (size_t)n * k * 8), H(c, (size_t)k * k * 8),
      W(c, (size_t)k * k * 8), Vw(c, syev_dc_work_doubles(k) * 8), G(c, (size_t)k * k * 8), Rinv(c, (size_t)k * k * 8),
      nrm(c, (size_t)k * 8), res(c, (size_t)(k + 2) * 8), info(c, (size_t)k * sizeof(int) + 16), scal(c, 16 * 8) {}

int topk_eig(Ctx* c, const double* A, int n, int want, int k, const double* X0, int maxit, double tol,
             double* X, double* AX, double* theta, EigStats* st) {
  hipStream_t s = c->stream;
  if (n <= JACOBI_MAX) {
    // full spectrum: exact reference semantics (all n singular values / vectors), by divide and conquer
    // (2-3.5x faster than Jacobi sweeps to convergence at n = 40..192)
    DevBuf Vw(c, syev_dc_work_doubles(n) * 8);
    syev_dc(A, n, n, theta, X, Vw.as<double>(), s);
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, n, n, n, n, n), A, X, AX, s);
    if (st) { st->iters = 1; st->k = n; st->resid = 0.0; }
    return n;
  }
  k = eig_block_width(n, want, k);
  // the entry points reject want > JACOBI_MAX for n > JACOBI_MAX (tnad_limits.h); the residual kernel and the
  // status transfer below are sized by want and need want <= k
  if (want < 1 || want > k) throw Error(TNAD_ERR_UNSUPPORTED, "eigensolver: kept block wider than 192");
  EigWork w(c, n, k);
  // starting block
  if (X0) {
    k_copy(X0, w.Y.as<double>(), (size_t)n * k, s);
  } else {
    k_random_matrix(w.Y.as<double>(), (size_t)n * k, 0x1234567u, s);
  }
  orthonormalize(c, w.Y.as<double>(), n, k, X, w, 0x9e3779b9u);
  std::vector<double> hres(k + 2), hth(k);
  // opt-in trace of the iteration (TNAD_DEBUG_EIG=1), to stderr
  static const bool dbg = [] {
    const char* e = std::getenv("TNAD_DEBUG_EIG");
    return e && *e && *e != '0';
  }();
  int it = 0;
  double rel = 0.0, best = HUGE_VAL, best_mark = HUGE_VAL, prev_rel = HUGE_VAL;
  int best_it = 0, n_exact = 0;
  bool force_exact = false;
  for (it = 1; it <= maxit; ++it) {
    // Y = A X ; H = X^T Y ; H = W diag(theta) W^T ; Xr = X W ; Yr = Y W
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, n, n, n, n), A, X, w.Y.as<double>(), s);
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', k, k, n, n, n, k), X, w.Y.as<double>(), w.H.as<double>(), s);
    dbg_finite(c, "X", X, (size_t)n * k);
    dbg_finite(c, "H", w.H.as<double>(), (size_t)k * k);
    // Rayleigh-Ritz: once the block is nearly an invariant subspace (H nearly diagonal in the
    // previous Ritz basis) a few rotations of near-degenerate pairs plus second-order perturbation
    // theory (ritz_select) stand in for the eigendecomposition of H; otherwise H is decomposed exactly,
    // by divide and conquer: on the first cold iteration, whenever H is not close enough to diagonal,
    // and whenever the residual stopped falling (by 10% per iteration).  ritz_select's flag gates the
    // exact solver's kernels on the device: no host round trip.  (Exact solves where two or three
    // Jacobi sweeps used to improve the basis cut the subspace iterations per evaluation from 35 to 30
    // on the FP64 path at chi = 128, D = 4, and 17 to 14 on the mixed one, 21% and 10% less time.)
    const bool exact_forced = force_exact || (it == 1 && !X0);
    int* gate = w.info.as<int>();
    if (!exact_forced) ritz_select(w.H.as<double>(), k, k, want, 1e-2, theta, w.W.as<double>(), gate, s);
    syev_dc(w.H.as<double>(), k, k, theta, w.W.as<double>(), w.Vw.as<double>(), s, exact_forced ? nullptr : gate);
    dbg_finite(c, "W", w.W.as<double>(), (size_t)k * k);
    dbg_finite(c, "theta", theta, (size_t)k);
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, k, n, k, n), X, w.W.as<double>(), w.Xr.as<double>(), s);
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, k, k, n, k, n), w.Y.as<double>(), w.W.as<double>(),
              w.Yr.as<double>(), s);
    hipLaunchKernelGGL(resid_norms_kernel, dim3(want), dim3(NT), 0, s, w.Xr.as<double>(), w.Yr.as<double>(),
                       theta, n, w.res.as<double>(), exact_forced ? nullptr : gate);
    TNAD_HIP_TRY(hipGetLastError());
    c->read_async(hres.data(), w.res.ptr, (size_t)(want + 2) * 8);   // residuals, theta_0, flag
    c->sync();
    hth[0] = hres[want];
    const int flag = (int)hres[want + 1];
    const bool exact = flag != 0;   // the exact solver ran (forced, or the selective Rayleigh-Ritz declined)
    n_exact += exact;
    double mx = 0.0;
    for (int j = 0; j < want; ++j) mx = std::max(mx, hres[j]);
    rel = mx / std::max(std::fabs(hth[0]), 1e-300);
    if (dbg)
      std::fprintf(stderr, "[tnad eig] n=%d k=%d want=%d it=%d rel_resid=%.3e theta0=%.6e rr=%s(%d)\n", n, k, want,
                   it, rel, hth[0], exact ? "exact" : "select", exact_forced ? -1 : flag);
    force_exact = !(rel < 0.9 * prev_rel);
    prev_rel = rel;
    if (!(rel == rel)) throw Error(TNAD_ERR_NAN, "eigensolver: NaN in Ritz residual");
    if (rel <= tol || it == maxit) break;
    // stagnation at the rounding level: no halving of the best residual in 8 iterations
    if (rel < best) { if (rel < 0.5 * best_mark) { best_mark = rel; best_it = it; } best = rel; }
    if (best < 1e3 * tol && it - best_it >= 8) break;
    hipLaunchKernelGGL(ritz_scale_norms_kernel, dim3(k), dim3(256), 0, s, w.Xr.as<double>(), w.Yr.as<double>(),
                       theta, n, w.Y.as<double>(), w.nrm.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
    orthonormalize(c, w.Y.as<double>(), n, k, X, w, 0x51ed2701u + it, true);
  }
  k_copy(w.Xr.as<double>(), X, (size_t)n * k, s);
  k_copy(w.Yr.as<double>(), AX, (size_t)n * k, s);
  hipLaunchKernelGGL(canon_sign_kernel, dim3(k), dim3(NT), 0, s, X, AX, n);
  TNAD_HIP_TRY(hipGetLastError());
  if (st) { st->iters = std::min(it, maxit); st->k = k; st->resid = rel; st->exact = n_exact; }
  return k;
}

// ==========================================================================
// CTMRG step (src/ctmrg.jl:126-153)
// ==========================================================================
StepTape::StepTape(Ctx* c, const StepDims& d, Prec p, int k)
    : prec(p), k(k), C(c, (size_t)d.chi * d.chi * 8), T(c, (size_t)d.chi * d.D * d.chi * 8),
      X(c, (size_t)d.n * k * 8), AX(c, (size_t)d.n * k * 8), theta(c, (size_t)k * 8),
      M(c, (size_t)d.n * d.n * 8), X2(c, p, (size_t)d.n * d.n), Y2(c, p, (size_t)d.n * d.n),
      C2(c, (size_t)d.chi * d.chi * 8), T2(c, (size_t)d.chi * d.D * d.chi * 8), nrm(c, 4 * 8) {}

void ctmrg_step(Ctx* c, const StepDims& d, const StepOpts& o, const double* a64, const PBuf& ap,
                const double* C, const double* T, const double* X0, double* X_keep, double* Cout, double* Tout,
                std::vector<double>& vals, StepTape* tape, StepState& state) {
  hipStream_t s = c->stream;
  const int chi = d.chi, D = d.D, n = d.n;
  const long long lc = chi, lD = D;
  const Prec p = o.prec;
  const size_t es = esize(p);
  // operands at contraction precision
  PBuf Cp(c, p, (size_t)chi * chi), Tp(c, p, (size_t)chi * D * chi);
  to_prec(c, C, Cp);
  to_prec(c, T, Tp);
  PBuf X1(c, p, (size_t)n * chi);
  PBuf X2tmp;
  PBuf* X2 = tape ? &tape->X2 : nullptr;
  if (!X2) { X2tmp = PBuf(c, p, (size_t)n * n); X2 = &X2tmp; }
  PBuf cp(c, p, (size_t)n * n);
  // cp = ein"ad,iba,dcl,jkcb -> ijlk"(C, T, T, a)          (src/ctmrg.jl:130)
  tc(c, p, cm("iba", {lc, lD, lc}), Tp.ptr(), cm("ad", {lc, lc}), Cp.ptr(), cm("ibd", {lc, lD, lc}), X1.ptr());
  tc(c, p, cm("ibd", {lc, lD, lc}), X1.ptr(), cm("dcl", {lc, lD, lc}), Tp.ptr(), cm("ibcl", {lc, lD, lD, lc}),
     X2->ptr());
  tc(c, p, cm("ibcl", {lc, lD, lD, lc}), X2->ptr(), cm("jkcb", {lD, lD, lD, lD}), ap.ptr(),
     cm("ijlk", {lc, lD, lc, lD}), cp.ptr());
  // cpmat = reshape(cp, n, n); cpmat += cpmat'                (src/ctmrg.jl:134-135)
  DevBuf Mtmp;
  double* M = tape ? tape->M.as<double>() : nullptr;
  if (!M) { Mtmp = DevBuf(c, (size_t)n * n * 8); M = Mtmp.as<double>(); }
  if (p == Prec::F64) {
    k_sym_add(cp.b.as<double>(), n, M, s);
  } else {
    DevBuf cp64(c, (size_t)n * n * 8);
    k_cast<float>(cp.b.as<float>(), cp64.as<double>(), (size_t)n * n, s);
    k_sym_add(cp64.as<double>(), n, M, s);
  }
  // u, s, v = svd(cpmat) -> top-k eigenpairs by |lambda|        (src/ctmrg.jl:136)
  // block size: n (full spectrum) or 2.25 chi guard-extended (eig_block if set), capped at 192
  const int kreq = step_k(d, o);
  DevBuf Xtmp, AXtmp, thtmp;
  double *X = nullptr, *AX = nullptr, *theta = nullptr;
  if (tape) { X = tape->X.as<double>(); AX = tape->AX.as<double>(); theta = tape->theta.as<double>(); }
  else {
    Xtmp = DevBuf(c, (size_t)n * kreq * 8); AXtmp = DevBuf(c, (size_t)n * kreq * 8); thtmp = DevBuf(c, kreq * 8);
    X = Xtmp.as<double>(); AX = AXtmp.as<double>(); theta = thtmp.as<double>();
  }
  EigStats est;
  const bool warm = X0 != nullptr && n > JACOBI_MAX;
  c->mark(4);
  int k = topk_eig(c, M, n, chi, kreq, warm ? X0 : nullptr, warm ? o.eig_warm_iters : o.eig_max_iters,
                   p == Prec::F64 ? o.eig_tol : o.eig_tol_low, X, AX, theta, &est);
  c->mark(5);
  state.k = k;
  if (X_keep) k_copy(X, X_keep, (size_t)n * k, s);
  state.eig_iters = est.iters;
  state.eig_resid = est.resid;
  // z = reshape(u[:, 1:chi], chi, D, chi)                        (src/ctmrg.jl:137)
  const double* z = X;   // first chi columns
  // C' + C'^T = z^T (cpmat + cpmat^T) z  (src/ctmrg.jl:139,145), FP64
  double* C2 = tape ? tape->C2.as<double>() : nullptr;
  DevBuf C2tmp;
  if (!C2) { C2tmp = DevBuf(c, (size_t)chi * chi * 8); C2 = C2tmp.as<double>(); }
  tc_launch(c, Prec::F64, gemm_desc('T', 'N', chi, chi, n, n, n, chi), z, AX, C2, s);
  // T' = ein"abjcd,abi,dck -> ijk"(tp, z, z), tp = ein"iam,jkla -> ijklm"(T, a)  (src/ctmrg.jl:131,140)
  // evaluated as ((T z) a) z without materialising tp.
  PBuf zp(c, p, (size_t)n * chi);
  to_prec(c, z, zp);
  PBuf Y1(c, p, (size_t)n * n);
  PBuf Y2tmp;
  PBuf* Y2 = tape ? &tape->Y2 : nullptr;
  if (!Y2) { Y2tmp = PBuf(c, p, (size_t)n * n); Y2 = &Y2tmp; }
  PBuf T1p(c, p, (size_t)chi * D * chi);
  tc(c, p, cm("aed", {lc, lD, lc}), Tp.ptr(), cm("abx", {lc, lD, lc}), zp.ptr(), cm("edbx", {lD, lc, lD, lc}),
     Y1.ptr());
  tc(c, p, cm("edbx", {lD, lc, lD, lc}), Y1.ptr(), cm("bjce", {lD, lD, lD, lD}), ap.ptr(),
     cm("dcxj", {lc, lD, lc, lD}), Y2->ptr());
  tc(c, p, cm("dcxj", {lc, lD, lc, lD}), Y2->ptr(), cm("dcy", {lc, lD, lc}), zp.ptr(), cm("xjy", {lc, lD, lc}),
     T1p.ptr());
  DevBuf T1(c, (size_t)chi * D * chi * 8);
  from_prec(c, T1p, T1.as<double>());
  double* T2 = tape ? tape->T2.as<double>() : nullptr;
  DevBuf T2tmp;
  if (!T2) { T2tmp = DevBuf(c, (size_t)chi * D * chi * 8); T2 = T2tmp.as<double>(); }
  // edge += ein"ijk -> kji"(conj(edge))                          (src/ctmrg.jl:146)
  k_edge_sym(T1.as<double>(), chi, D, T2, s);
  // corner /= norm(corner); edge /= norm(edge)                   (src/ctmrg.jl:149-150)
  DevBuf red(c, 300 * 8), nr(c, 4 * 8);
  double* nrm = tape ? tape->nrm.as<double>() : nr.as<double>();
  k_sumsq(C2, (size_t)chi * chi, red.as<double>(), s);
  k_copy(red.as<double>(), nrm + 0, 1, s);
  k_sumsq(T2, (size_t)chi * D * chi, red.as<double>(), s);
  k_copy(red.as<double>(), nrm + 1, 1, s);
  hipLaunchKernelGGL(sqrt_inplace_kernel, dim3(1), dim3(1), 0, s, nrm + 0);
  hipLaunchKernelGGL(sqrt_inplace_kernel, dim3(1), dim3(1), 0, s, nrm + 1);
  k_copy(C2, Cout, (size_t)chi * chi, s);
  k_scale_by_dev(Cout, (size_t)chi * chi, nrm + 0, s);
  k_copy(T2, Tout, (size_t)chi * D * chi, s);
  k_scale_by_dev(Tout, (size_t)chi * D * chi, nrm + 1, s);
  if (tape) {
    tape->eig_tol = p == Prec::F64 ? o.eig_tol : o.eig_tol_low;
    k_copy(C, tape->C.as<double>(), (size_t)chi * chi, s);
    k_copy(T, tape->T.as<double>(), (size_t)chi * D * chi, s);
  }
  // vals = s ./ s[1]   (src/ctmrg.jl:142).  With the whole spectrum computed (n <= 192) all n
  // values; otherwise the k largest (the chi kept values, resolved to the eigensolver tolerance, and the
  // guard block's Ritz values, which are not), with the rest 0.  For n > 192 the step also measures the
  // norm of the spectrum beyond the kept chi values, rest = |M - Z Theta Z^T|_F / s[1] (Z the kept
  // eigenvectors): formed elementwise, it is accurate to ~n eps |M|_F even where rest is far below
  // sqrt(eps) (|M|_F^2 - sum theta^2 would lose it to cancellation).  The stop test uses it as a lower
  // bound on the change of the discarded values (run_ctmrg).
  const int nv = k;
  std::vector<double> th(nv);
  double sq[2] = {0.0, 0.0};
  DevBuf E, XT, red1, red2, sqd;
  if (n > JACOBI_MAX) {
    XT = DevBuf(c, (size_t)n * chi * 8);
    E = DevBuf(c, (size_t)n * n * 8);
    red1 = DevBuf(c, 300 * 8);
    red2 = DevBuf(c, 300 * 8);
    sqd = DevBuf(c, 2 * 8);
    k_copy(X, XT.as<double>(), (size_t)n * chi, s);
    k_scale_cols(XT.as<double>(), n, chi, theta, 0, s);
    k_copy(M, E.as<double>(), (size_t)n * n, s);
    tc_launch(c, Prec::F64, gemm_desc('N', 'T', n, n, chi, n, n, n, -1.0, 1.0), XT.as<double>(), X,
              E.as<double>(), s);
    k_sumsq(E.as<double>(), (size_t)n * n, red1.as<double>(), s);
    k_sumsq(M, (size_t)n * n, red2.as<double>(), s);
    k_copy(red1.as<double>(), sqd.as<double>() + 0, 1, s);
    k_copy(red2.as<double>(), sqd.as<double>() + 1, 1, s);
    c->read_async(sq, sqd.ptr, 2 * 8);
  }
  c->read_async(th.data(), theta, (size_t)nv * 8);
  c->sync();
  state.eig_ms = c->elapsed(4, 5);
  vals.assign(n, 0.0);
  double s0 = std::fabs(th[0]);
  for (int i = 0; i < nv; ++i) vals[i] = std::fabs(th[i]) / s0;
  for (int i = 0; i < nv; ++i)
    if (!(vals[i] == vals[i])) throw Error(TNAD_ERR_NAN, "ctmrgstep: NaN in singular values");
  if (n > JACOBI_MAX) {
    state.rest = std::sqrt(sq[0]) / s0;
    // rounding of the n^2 entries of E and of its sum (~n eps |M|_F), plus the kept residuals R = M Z - Z Theta:
    // |E - E_exact|_F <= sqrt(2) |R|_F <= sqrt(2 chi) max_j |r_j| (Mirsky)
    state.rest_err = (double)n * 2.3e-16 * std::sqrt(sq[1]) / s0 + std::sqrt(2.0 * chi) * est.resid;
    if (!(state.rest == state.rest)) throw Error(TNAD_ERR_NAN, "ctmrgstep: NaN in the discarded spectrum");
  }
  (void)es;
}

// ==========================================================================
// Eigen adjoint: cotangent of z = U[:, 1:chi] (plus the C' chain's C1bar)
// -> cotangent of cpmat, i.e. svd_back (src/trg.jl:72-106) with dU nonzero in
// the first chi columns, dS = dV = nothing, followed by cpmat += cpmat'.
//
//   cpbar = W U1^T + U1 W^T,  W = U1 Kt/2 + (X_c Lambda1 + A X_c)
//
// Kt/2 holds the pairs among the kept chi eigenvectors (reference formula
// verbatim).  Every other eigenpair i contributes g_ij / (lambda_j - lambda_i)
// along u_i, i.e. (lambda_j - A)^-1 = (lambda_j + A)(lambda_j^2 - A^2)^-1 on the
// complement of the kept vectors: X_c solves (lambda_j^2 - A^2) x_j = P G_j there
// (resolvent_pcg, preconditioned by the extra Ritz pairs the eigensolver carries).
// With the whole spectrum known (n <= 192) the preconditioner is the exact
// inverse and the solve ends after its first step.
// ==========================================================================
void eig_adjoint(Ctx* c, const StepDims& d, const StepTape& tp, const double* G, const double* C1bar,
                 double* cpbar, EigBackStats* st, double* sweep_ref) {
  hipStream_t s = c->stream;
  const int chi = d.chi, n = d.n, k = tp.k;
  const double* X = tp.X.as<double>();
  const double* lam = tp.theta.as<double>();
  DevBuf g(c, (size_t)chi * chi * 8), coef(c, (size_t)chi * chi * 8), W(c, (size_t)n * chi * 8);
  // g = U1^T G  (chi x chi)
  tc_launch(c, Prec::F64, gemm_desc('T', 'N', chi, chi, n, n, n, chi), X, G, g.as<double>(), s);
  hipLaunchKernelGGL(eigback_coef_kernel, dim3((chi * chi + NT - 1) / NT), dim3(NT), 0, s, g.as<double>(), lam,
                     C1bar, chi, 1e-40, coef.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  // W = U1 coef  (n x chi . chi x chi)
  tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, chi, chi, n, chi, n), X, coef.as<double>(), W.as<double>(), s);
  int iters = 0;
  if (n > chi) {
    DevBuf sq(c, (size_t)k * 8), Xc(c, (size_t)n * chi * 8), AXc(c, (size_t)n * chi * 8),
        tmp(c, (size_t)n * chi * 8);
    k_square(lam, sq.as<double>(), k, s);
    const double* M = tp.M.as<double>();
    auto applyA2 = [&](const double* V, double* out) {
      tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, chi, n, n, n, n), M, V, tmp.as<double>(), s);
      tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, chi, n, n, n, n), M, tmp.as<double>(), out, s);
    };
    // A step computed with low-precision contractions resolved its eigenpairs only to the relative
    // residual its contractions can support (1e-8 and above), so this adjoint is no more accurate
    // than that whatever the solve does: its complement part is solved to 1e-6 instead of 1e-13, and
    // no further than 1e-13 of the reverse sweep's largest right-hand side (the early steps of a
    // converging CTMRG receive cotangents many orders smaller than the last ones, so that bound is the
    // one that holds their error down).  The same holds for an FP64 step whose eigensolver ran to a
    // looser residual than 1e-13 (the mixed path ties it to the step-to-step change).  At chi = 128,
    // D = 4 the relative bar at 1e-6 instead of 1e-8 cut the reverse sweep from 11.2 to 10.2 ms with the
    // gradient unchanged (relative difference to the FP64 path 4.8e-9 either way).
    const bool low = tp.prec != Prec::F64;
    const bool loose = low || tp.eig_tol > 1e-13;
    const double pcg_tol = loose ? 1e-6 : 1e-13;
    iters = resolvent_pcg(c, n, chi, k, X, sq.as<double>(), sq.as<double>() + chi, applyA2, G, Xc.as<double>(),
                          pcg_tol, 200, 1e-40, sweep_ref, loose ? 1e-13 : 0.0);
    static const bool dbg = [] {
      const char* e = std::getenv("TNAD_DEBUG_EIG");
      return e && *e && *e != '0';
    }();
    if (dbg)
      std::fprintf(stderr, "[tnad eig] adjoint n=%d chi=%d prec=%s eig_tol=%.1e: %d resolvent CG iterations (tol %.1e)\n",
                   n, chi, low ? "low" : "fp64", tp.eig_tol, iters, pcg_tol);
    // W += X_c Lambda1 + A X_c
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, chi, n, n, n, n), M, Xc.as<double>(), AXc.as<double>(), s);
    hipLaunchKernelGGL(add_lam_cols_kernel, dim3(nb((size_t)n * chi)), dim3(NT), 0, s, W.as<double>(),
                       Xc.as<double>(), AXc.as<double>(), lam, n, chi);
    TNAD_HIP_TRY(hipGetLastError());
  }
  // cpbar = W U1^T + U1 W^T
  tc_launch(c, Prec::F64, gemm_desc('N', 'T', n, n, chi, n, n, n), W.as<double>(), X, cpbar, s);
  tc_launch(c, Prec::F64, gemm_desc('N', 'T', n, n, chi, n, n, n, 1.0, 1.0), X, W.as<double>(), cpbar, s);
  if (st) st->pcg_iters = iters;
}

// ==========================================================================
// Reverse sweep of one ctmrgstep.  In: cotangents of the step outputs
// (C3bar, T3bar).  Out: cotangents of the step inputs (Cbar, Tbar) and the
// bulk contribution added into abar.
// ==========================================================================
void ctmrg_step_back(Ctx* c, const StepDims& d, Prec p, const StepTape& tp, const double* a64, const PBuf& ap,
                     const double* C3bar, const double* T3bar, double* Cbar, double* Tbar, double* abar,
                     double* sweep_ref) {
  hipStream_t s = c->stream;
  const int chi = d.chi, D = d.D, n = d.n;
  const long long lc = chi, lD = D;
  const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi, nA = (size_t)D * D * D * D;
  DevBuf red(c, 300 * 8), dotb(c, 300 * 8);
  // --- normalisation adjoints (src/autodiff.jl:23-29) --------------------
  DevBuf C2bar(c, nC * 8), T2bar(c, nT * 8), C1bar(c, nC * 8), T1bar(c, nT * 8);
  {
    DevBuf ss(c, 2 * 8), dd(c, 2 * 8);
    k_sumsq(tp.C2.as<double>(), nC, red.as<double>(), s);
    k_copy(red.as<double>(), ss.as<double>() + 0, 1, s);
    k_dot(tp.C2.as<double>(), C3bar, nC, red.as<double>(), s);
    k_copy(red.as<double>(), dd.as<double>() + 0, 1, s);
    k_sumsq(tp.T2.as<double>(), nT, red.as<double>(), s);
    k_copy(red.as<double>(), ss.as<double>() + 1, 1, s);
    k_dot(tp.T2.as<double>(), T3bar, nT, red.as<double>(), s);
    k_copy(red.as<double>(), dd.as<double>() + 1, 1, s);
    hipLaunchKernelGGL(norm_back_kernel, dim3(nb(nC)), dim3(NT), 0, s, tp.C2.as<double>(), C3bar,
                       ss.as<double>() + 0, dd.as<double>() + 0, C2bar.as<double>(), nC);
    hipLaunchKernelGGL(norm_back_kernel, dim3(nb(nT)), dim3(NT), 0, s, tp.T2.as<double>(), T3bar,
                       ss.as<double>() + 1, dd.as<double>() + 1, T2bar.as<double>(), nT);
    TNAD_HIP_TRY(hipGetLastError());
  }
  // C1bar = C2bar + C2bar' ; T1bar = T2bar + permutedims(T2bar, (3,2,1))
  k_sym_add(C2bar.as<double>(), chi, C1bar.as<double>(), s);
  k_edge_sym(T2bar.as<double>(), chi, D, T1bar.as<double>(), s);

  // tape intermediates stored at the forward step's precision, converted to the sweep's
  PBuf Y2tmp, X2tmp;
  const PBuf& Y2 = at_prec(c, tp.Y2, p, Y2tmp);
  const PBuf& X2 = at_prec(c, tp.X2, p, X2tmp);
  // --- T' chain adjoint (tp never materialised) ----------------------------
  PBuf T1b(c, p, nT), zp(c, p, (size_t)n * chi), Tp(c, p, nT), Cp(c, p, nC);
  to_prec(c, T1bar.as<double>(), T1b);
  to_prec(c, tp.X.as<double>(), zp);   // first n*chi entries = z
  to_prec(c, tp.T.as<double>(), Tp);
  to_prec(c, tp.C.as<double>(), Cp);
  PBuf P1(c, p, (size_t)n * n), Q1(c, p, (size_t)n * n), S1(c, p, (size_t)n * n);
  PBuf Tbp(c, p, nT), abp(c, p, nA), zbp(c, p, (size_t)n * chi);
  // P1[a,b,j,y] = sum_x z[a,b,x] T1bar[x,j,y]
  tc(c, p, cm("abx", {lc, lD, lc}), zp.ptr(), cm("xjy", {lc, lD, lc}), T1b.ptr(), cm("abjy", {lc, lD, lD, lc}),
     P1.ptr());
  // Q1[a,e,c,y] = sum_{b,j} P1[a,b,j,y] a[b,j,c,e] ; Tbar[a,e,d] = sum_{c,y} Q1[a,e,c,y] z[d,c,y]
  tc(c, p, cm("abjy", {lc, lD, lD, lc}), P1.ptr(), cm("bjce", {lD, lD, lD, lD}), ap.ptr(),
     cm("aecy", {lc, lD, lD, lc}), Q1.ptr());
  tc(c, p, cm("aecy", {lc, lD, lD, lc}), Q1.ptr(), cm("dcy", {lc, lD, lc}), zp.ptr(), cm("aed", {lc, lD, lc}),
     Tbp.ptr());
  // S1[a,e,c,y] = sum_d T[a,e,d] z[d,c,y] ; abar[b,j,c,e] = sum_{a,y} P1[a,b,j,y] S1[a,e,c,y]
  tc(c, p, cm("aed", {lc, lD, lc}), Tp.ptr(), cm("dcy", {lc, lD, lc}), zp.ptr(), cm("aecy", {lc, lD, lD, lc}),
     S1.ptr());
  tc(c, p, cm("abjy", {lc, lD, lD, lc}), P1.ptr(), cm("aecy", {lc, lD, lD, lc}), S1.ptr(),
     cm("bjce", {lD, lD, lD, lD}), abp.ptr());
  // U1[a,b,j,y] = sum_{e,c} S1[a,e,c,y] a[b,j,c,e] (reuses P1) ; zbar[a,b,x] = sum_{j,y} U1 T1bar[x,j,y]
  tc(c, p, cm("aecy", {lc, lD, lD, lc}), S1.ptr(), cm("bjce", {lD, lD, lD, lD}), ap.ptr(),
     cm("abjy", {lc, lD, lD, lc}), P1.ptr());
  tc(c, p, cm("abjy", {lc, lD, lD, lc}), P1.ptr(), cm("xjy", {lc, lD, lc}), T1b.ptr(), cm("abx", {lc, lD, lc}),
     zbp.ptr());
  // zbar[d,c,y] += sum_{x,j} Y2[d,c,x,j] T1bar[x,j,y]
  tc(c, p, cm("dcxj", {lc, lD, lc, lD}), Y2.ptr(), cm("xjy", {lc, lD, lc}), T1b.ptr(), cm("dcy", {lc, lD, lc}),
     zbp.ptr(), 1.0, 1.0);
  DevBuf zbar(c, (size_t)n * chi * 8);
  from_prec(c, zbp, zbar.as<double>());
  DevBuf a_acc(c, nA * 8);
  from_prec(c, abp, a_acc.as<double>());
  k_axpby(1.0, a_acc.as<double>(), 1.0, abar, nA, s);
  DevBuf Tacc(c, nT * 8);
  from_prec(c, Tbp, Tacc.as<double>());

  // --- C' chain: zbar += (cpmat + cpmat') z C1bar  (FP64; A z kept from the forward)
  tc_launch(c, Prec::F64, gemm_desc('N', 'N', n, chi, chi, n, chi, n, 1.0, 1.0), tp.AX.as<double>(),
            C1bar.as<double>(), zbar.as<double>(), s);

  // --- eigen adjoint -> cpbar (n x n, FP64) ---------------------------------
  DevBuf cpbar(c, (size_t)n * n * 8);
  eig_adjoint(c, d, tp, zbar.as<double>(), C1bar.as<double>(), cpbar.as<double>(), nullptr, sweep_ref);

  // --- cp chain adjoint -----------------------------------------------------
  PBuf cpb(c, p, (size_t)n * n), X2b(c, p, (size_t)n * n), X1(c, p, (size_t)n * chi), X1b(c, p, (size_t)n * chi);
  PBuf Cbp(c, p, nC), Tbp2(c, p, nT);
  to_prec(c, cpbar.as<double>(), cpb);
  // X1 = T C (recomputed)
  tc(c, p, cm("iba", {lc, lD, lc}), Tp.ptr(), cm("ad", {lc, lc}), Cp.ptr(), cm("ibd", {lc, lD, lc}), X1.ptr());
  // X2bar[i,b,c,l] = sum_{j,k} cpbar[i,j,l,k] a[j,k,c,b]
  tc(c, p, cm("ijlk", {lc, lD, lc, lD}), cpb.ptr(), cm("jkcb", {lD, lD, lD, lD}), ap.ptr(),
     cm("ibcl", {lc, lD, lD, lc}), X2b.ptr());
  // abar[j,k,c,b] += sum_{i,l} cpbar[i,j,l,k] X2[i,b,c,l]
  tc(c, p, cm("ijlk", {lc, lD, lc, lD}), cpb.ptr(), cm("ibcl", {lc, lD, lD, lc}), X2.ptr(),
     cm("jkcb", {lD, lD, lD, lD}), abp.ptr());
  from_prec(c, abp, a_acc.as<double>());
  k_axpby(1.0, a_acc.as<double>(), 1.0, abar, nA, s);
  // X1bar[i,b,d] = sum_{c,l} X2bar[i,b,c,l] T[d,c,l] ; Tbar[d,c,l] += sum_{i,b} X1[i,b,d] X2bar[i,b,c,l]
  tc(c, p, cm("ibcl", {lc, lD, lD, lc}), X2b.ptr(), cm("dcl", {lc, lD, lc}), Tp.ptr(), cm("ibd", {lc, lD, lc}),
     X1b.ptr());
  tc(c, p, cm("ibd", {lc, lD, lc}), X1.ptr(), cm("ibcl", {lc, lD, lD, lc}), X2b.ptr(), cm("dcl", {lc, lD, lc}),
     Tbp2.ptr());
  // Cbar[a,d] = sum_{i,b} T[i,b,a] X1bar[i,b,d] ; Tbar[i,b,a] += sum_d X1bar[i,b,d] C[a,d]
  tc(c, p, cm("iba", {lc, lD, lc}), Tp.ptr(), cm("ibd", {lc, lD, lc}), X1b.ptr(), cm("ad", {lc, lc}), Cbp.ptr());
  tc(c, p, cm("ibd", {lc, lD, lc}), X1b.ptr(), cm("ad", {lc, lc}), Cp.ptr(), cm("iba", {lc, lD, lc}), Tbp2.ptr(),
     1.0, 1.0);
  from_prec(c, Cbp, Cbar);
  from_prec(c, Tbp2, Tbar);
  k_axpby(1.0, Tacc.as<double>(), 1.0, Tbar, nT, s);
  (void)a64;
}

// ==========================================================================
// expectationvalue (src/variationalipeps.jl:49-56), FP64, with reverse sweep
// l = ein"ab,ica,bde,cjfdlm,eg,gfk -> ijklm"(C, T, T, ap, C, T)
// ==========================================================================
EvCache::EvCache(Ctx* c, int chi, int D, int ph)
    : apn(c, (size_t)D * D * D * D * ph * ph * 8), E1(c, (size_t)chi * D * chi * 8),
      E2(c, (size_t)chi * D * D * chi * 8), E3(c, (size_t)chi * D * chi * 8),
      E4(c, (size_t)chi * D * chi * D * ph * ph * 8), l(c, (size_t)chi * D * chi * ph * ph * 8),
      tr(c, (size_t)chi * D * chi * 8), scal(c, 8 * 8) {}

void expectation_fwd(Ctx* c, int chi, int D, int ph, const double* h, const double* ap, const double* C,
                     const double* T, EvCache& ev, double* e_host) {
  hipStream_t s = c->stream;
  const long long lc = chi, lD = D, ls = ph;
  const long long n = lc * lD;
  const size_t nap = (size_t)D * D * D * D * ph * ph;
  DevBuf red(c, 300 * 8);
  // ap /= norm(ap)
  k_sumsq(ap, nap, red.as<double>(), s);
  k_copy(red.as<double>(), ev.scal.as<double>() + 3, 1, s);   // |ap|^2
  k_copy(ap, ev.apn.as<double>(), nap, s);
  k_copy(red.as<double>(), ev.scal.as<double>() + 4, 1, s);
  hipLaunchKernelGGL(sqrt_inplace_kernel, dim3(1), dim3(1), 0, s, ev.scal.as<double>() + 4);
  k_scale_by_dev(ev.apn.as<double>(), nap, ev.scal.as<double>() + 4, s);
  const Prec F = Prec::F64;
  double* E1 = ev.E1.as<double>();
  double* E2 = ev.E2.as<double>();
  double* E3 = ev.E3.as<double>();
  double* E4 = ev.E4.as<double>();
  double* l = ev.l.as<double>();
  tc(c, F, cm("ica", {lc, lD, lc}), T, cm("ab", {lc, lc}), C, cm("icb", {lc, lD, lc}), E1);
  tc(c, F, cm("icb", {lc, lD, lc}), E1, cm("bde", {lc, lD, lc}), T, cm("icde", {lc, lD, lD, lc}), E2);
  tc(c, F, cm("eg", {lc, lc}), C, cm("gfk", {lc, lD, lc}), T, cm("efk", {lc, lD, lc}), E3);
  // E4[i,j,e,f,l,m] = sum_{c,d} E2[i,c,d,e] ap[c,j,f,d,l,m]  (one launch per physical pair (l,m))
  const long long e4s[4] = {1, lc, n, n * lc};
  for (int lm = 0; lm < ph * ph; ++lm) {
    tc(c, F, cm("icde", {lc, lD, lD, lc}), E2, cm("cjfd", {lD, lD, lD, lD}), ev.apn.as<double>() + (size_t)lm * D * D * D * D,
       with_strides("ijef", {lc, lD, lc, lD}, {e4s[0], e4s[1], e4s[2], e4s[3]}), E4 + (size_t)lm * n * n);
  }
  // l[i,j,k,l,m] = sum_{e,f} E4[i,j,e,f,l,m] E3[e,f,k]
  tc(c, F, with_strides("ijeflm", {lc, lD, lc, lD, ls, ls}, {1, lc, n, n * lc, n * n, n * n * ls}), E4,
     cm("efk", {lc, lD, lc}), E3, cm("ijklm", {lc, lD, lc, ls, ls}), l);
  // e = ein"abcij,abckl,ijkl -> "(l, l, h) ; n = ein"ijkaa,ijkbb -> "(l, l)
  DevBuf Lh(c, (size_t)n * chi * ph * ph * 8);
  tc(c, F, cm("abcij", {lc, lD, lc, ls, ls}), l, cm("ijkl", {ls, ls, ls, ls}), h, cm("abckl", {lc, lD, lc, ls, ls}),
     Lh.as<double>());
  const size_t nl = (size_t)n * chi * ph * ph, m3 = (size_t)n * chi;
  k_dot(Lh.as<double>(), l, nl, red.as<double>(), s);
  k_copy(red.as<double>(), ev.scal.as<double>() + 0, 1, s);   // e
  hipLaunchKernelGGL(trace_last2_kernel, dim3((m3 + NT - 1) / NT), dim3(NT), 0, s, l, (int)m3, ph, ev.tr.as<double>());
  k_sumsq(ev.tr.as<double>(), m3, red.as<double>(), s);
  k_copy(red.as<double>(), ev.scal.as<double>() + 1, 1, s);   // nn
  // scal[5] = e/nn, scal[6] = 1/nn, scal[7] = -e/nn^2
  hipLaunchKernelGGL(energy_scalars_kernel, dim3(1), dim3(1), 0, s, ev.scal.as<double>() + 0,
                     ev.scal.as<double>() + 1, ev.scal.as<double>() + 5);
  TNAD_HIP_TRY(hipGetLastError());
  if (e_host) {
    c->read_async(e_host, ev.scal.as<double>() + 5, 8);
    c->sync();
  }
}

void expectation_back(Ctx* c, int chi, int D, int ph, const double* h, const double* ap, const double* C,
                      const double* T, EvCache& ev, double* apbar, double* Cbar, double* Tbar) {
  hipStream_t s = c->stream;
  const long long lc = chi, lD = D, ls = ph;
  const long long n = lc * lD;
  const size_t nap = (size_t)D * D * D * D * ph * ph;
  const size_t nl = (size_t)n * chi * ph * ph, m3 = (size_t)n * chi;
  const Prec F = Prec::F64;
  double* l = ev.l.as<double>();
  // lbar = ebar (l h^T + l h) + 2 nnbar tr (x) delta
  DevBuf lbar(c, nl * 8), tmp(c, nl * 8), red(c, 300 * 8);
  tc(c, F, cm("abcij", {lc, lD, lc, ls, ls}), l, cm("ijkl", {ls, ls, ls, ls}), h, cm("abckl", {lc, lD, lc, ls, ls}),
     lbar.as<double>());
  tc(c, F, cm("abckl", {lc, lD, lc, ls, ls}), l, cm("ijkl", {ls, ls, ls, ls}), h, cm("abcij", {lc, lD, lc, ls, ls}),
     tmp.as<double>());
  k_axpby(1.0, tmp.as<double>(), 1.0, lbar.as<double>(), nl, s);
  // scale by ebar (device scalar): lbar *= 1/nn  -> use scale_by_dev with nn
  k_scale_by_dev(lbar.as<double>(), nl, ev.scal.as<double>() + 1, s);
  {
    // 2 * nnbar = -2 e / nn^2
    DevBuf cf(c, 8);
    std::vector<double> sc(8);
    c->read_async(sc.data(), ev.scal.ptr, 8 * 8);
    c->sync();
    double coef = 2.0 * sc[7];
    hipLaunchKernelGGL(add_trace_diag_kernel, dim3((m3 * ph + NT - 1) / NT), dim3(NT), 0, s, lbar.as<double>(),
                       ev.tr.as<double>(), (const double*)nullptr, coef, (int)m3, ph);
    TNAD_HIP_TRY(hipGetLastError());
  }
  double* E1 = ev.E1.as<double>();
  double* E2 = ev.E2.as<double>();
  double* E3 = ev.E3.as<double>();
  double* E4 = ev.E4.as<double>();
  const TView vE4 = with_strides("ijeflm", {lc, lD, lc, lD, ls, ls}, {1, lc, n, n * lc, n * n, n * n * ls});
  DevBuf E4b(c, (size_t)n * n * ph * ph * 8), E3b(c, (size_t)n * chi * 8), E2b(c, (size_t)n * n * 8),
      E1b(c, (size_t)n * chi * 8), apnb(c, nap * 8);
  tc(c, F, cm("ijklm", {lc, lD, lc, ls, ls}), lbar.as<double>(), cm("efk", {lc, lD, lc}), E3, vE4, E4b.as<double>());
  tc(c, F, vE4, E4, cm("ijklm", {lc, lD, lc, ls, ls}), lbar.as<double>(), cm("efk", {lc, lD, lc}), E3b.as<double>());
  const TView vE4s = with_strides("ijef", {lc, lD, lc, lD}, {1, lc, n, n * lc});
  for (int lm = 0; lm < ph * ph; ++lm) {
    const size_t oa = (size_t)lm * D * D * D * D, oe = (size_t)lm * n * n;
    tc(c, F, vE4s, E4b.as<double>() + oe, cm("cjfd", {lD, lD, lD, lD}), ev.apn.as<double>() + oa,
       cm("icde", {lc, lD, lD, lc}), E2b.as<double>(), 1.0, lm == 0 ? 0.0 : 1.0);
    tc(c, F, cm("icde", {lc, lD, lD, lc}), E2, vE4s, E4b.as<double>() + oe, cm("cjfd", {lD, lD, lD, lD}),
       apnb.as<double>() + oa);
  }
  tc(c, F, cm("icde", {lc, lD, lD, lc}), E2b.as<double>(), cm("bde", {lc, lD, lc}), T, cm("icb", {lc, lD, lc}),
     E1b.as<double>());
  tc(c, F, cm("icb", {lc, lD, lc}), E1, cm("icde", {lc, lD, lD, lc}), E2b.as<double>(), cm("bde", {lc, lD, lc}),
     Tbar, 1.0, 1.0);
  tc(c, F, cm("icb", {lc, lD, lc}), E1b.as<double>(), cm("ab", {lc, lc}), C, cm("ica", {lc, lD, lc}), Tbar, 1.0, 1.0);
  tc(c, F, cm("ica", {lc, lD, lc}), T, cm("icb", {lc, lD, lc}), E1b.as<double>(), cm("ab", {lc, lc}), Cbar, 1.0, 1.0);
  tc(c, F, cm("efk", {lc, lD, lc}), E3b.as<double>(), cm("gfk", {lc, lD, lc}), T, cm("eg", {lc, lc}), Cbar, 1.0, 1.0);
  tc(c, F, cm("eg", {lc, lc}), C, cm("efk", {lc, lD, lc}), E3b.as<double>(), cm("gfk", {lc, lD, lc}), Tbar, 1.0, 1.0);
  // ap /= norm(ap) adjoint
  DevBuf dd(c, 8);
  k_dot(ap, apnb.as<double>(), nap, red.as<double>(), s);
  k_copy(red.as<double>(), dd.as<double>(), 1, s);
  hipLaunchKernelGGL(norm_back_kernel, dim3(nb(nap)), dim3(NT), 0, s, ap, apnb.as<double>(),
                     ev.scal.as<double>() + 3, dd.as<double>(), apbar, nap);
  TNAD_HIP_TRY(hipGetLastError());
}

}  // namespace tnad
