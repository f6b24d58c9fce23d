// Preconditioned conjugate gradients, one independent system per column, for the
// part of the eigen / singular-value adjoint that lies outside the kept vectors.
//
// svd_back (src/trg.jl:72-106) sums over every singular pair; a truncated
// eigensolver knows only the top k.  The sum over the pairs outside the kept r
// equals a solve with the resolvent of G = A^2 (CTMRG corner matrix) or M M^T /
// M^T M (TRG split) on the orthogonal complement of the kept vectors; the extra
// Ritz vectors the eigensolver carries precondition that solve.  Every block
// operation is a launch on the context stream; the host reads the per-column scalars
// once per iteration to stop.
#include "resolvent.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tnad {
namespace {

constexpr int NT = 256;

inline unsigned nb(size_t n) {
  size_t b = (n + NT - 1) / NT;
  if (b > 65535ull * 8) b = 65535ull * 8;
  return (unsigned)(b ? b : 1);
}

__global__ void square_kernel(const double* __restrict__ in, double* __restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * in[i];
}

// out[j] = sum_i X[i, j] Y[i, j]   (one block per column)
__global__ void coldot_kernel(const double* __restrict__ X, const double* __restrict__ Y, int rows,
                              double* __restrict__ out) {
  __shared__ double red[NT / 64];
  const int j = blockIdx.x;
  const double* x = X + (size_t)rows * j;
  const double* y = Y + (size_t)rows * j;
  double acc = 0.0;
  for (int i = threadIdx.x; i < rows; i += NT) acc += x[i] * y[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += red[w];
    out[j] = t;
  }
}

// isig_j = 1 / sig_j, or 0 for a kept value at the rounding level of the largest
// (sig_j <= 1e-28 sig_0, i.e. s_j <= 1e-14 s_0): such a column has no resolvable
// complement (the reference's Lorentzian gives F = 0 between equal zeros) and is skipped.
__global__ void inv_sig_kernel(const double* __restrict__ sig, int r, double* __restrict__ isig) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= r) return;
  const double s0 = sig[0];
  isig[j] = sig[j] > 1e-28 * s0 ? 1.0 / sig[j] : 0.0;
}

// One launch, two index ranges.  The first k r elements: preconditioner coefficients Cf from
// P = W^T R (k x r):
//   rows i <  r : -P_ij / sig_j                      (remove the kept directions)
//   rows i >= r :  P_ij (L(sig_j - tz_i) - 1/sig_j)  (tail directions at their own shift)
// evaluated without FP contraction so that an exactly degenerate pair gives L = 0.  The next rows r:
// out[:, j] = R[:, j] isig_j.
__global__ void precond_coef_div_kernel(const double* __restrict__ P, const double* __restrict__ sig,
                                        const double* __restrict__ isig, const double* __restrict__ tz, int r,
                                        int k, double eta, double* __restrict__ Cf, const double* __restrict__ R,
                                        int rows, double* __restrict__ out) {
#pragma clang fp contract(off)
  const size_t ncf = (size_t)k * r, tot = ncf + (size_t)rows * r;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    if (e < ncf) {
      const int i = (int)(e % k), j = (int)(e / k);
      const double inv = isig[j];
      double v = 0.0;
      if (inv != 0.0) {
        if (i < r) {
          v = -P[e] * inv;
        } else {
          double d = sig[j] - tz[i - r];
          double L = d / (d * d + eta);
          v = P[e] * (L - inv);
        }
      }
      Cf[e] = v;
    } else {
      const size_t f = e - ncf;
      out[f] = R[f] * isig[f / rows];
    }
  }
}

// Q[:, j] = V[:, j] sig_j - GV[:, j]
__global__ void shift_sub_kernel(const double* __restrict__ V, const double* __restrict__ GV,
                                 const double* __restrict__ sig, int rows, int r, double* __restrict__ Q) {
  size_t tot = (size_t)rows * r;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x)
    Q[e] = V[e] * sig[e / rows] - GV[e];
}

// x[j] *= y[j]
__global__ void mul_kernel(double* __restrict__ x, const double* __restrict__ y, int r) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < r) x[j] *= y[j];
}

// sum_i X[i, j] Y[i, j] over column j by one block of NT threads (the order of coldot_kernel), returned
// to every thread
__device__ __forceinline__ double block_coldot(const double* __restrict__ x, const double* __restrict__ y, int rows,
                                               double* red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < rows; i += NT) acc += x[i] * y[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < NT / 64; ++w) t += red[w];
  return t;
}

// One block per column j:  pq_j = p_j . q_j ; alpha_j = active_j && pq_j > 0 ? rz_j / pq_j : 0 (a
// column with p.Ap <= 0 is frozen) ; x_j += alpha_j p_j ; r_j -= alpha_j q_j
__global__ void cg_step_xr_kernel(double* __restrict__ x, double* __restrict__ res, const double* __restrict__ p,
                                  const double* __restrict__ q, const double* __restrict__ rz,
                                  double* __restrict__ active, int rows) {
  __shared__ double red[NT / 64];
  const int j = blockIdx.x;
  const size_t o = (size_t)rows * j;
  const double pq = block_coldot(p + o, q + o, rows, red);
  const bool ok = active[j] != 0.0 && pq > 0.0;
  const double a = ok ? rz[j] / pq : 0.0;
  if (!ok) {
    if (threadIdx.x == 0) active[j] = 0.0;
    return;
  }
  for (int i = threadIdx.x; i < rows; i += NT) {
    x[o + i] += a * p[o + i];
    res[o + i] -= a * q[o + i];
  }
}

// One block per column j:  rz_new = r_j . z_j ; beta_j = active_j && rz_j != 0 ? rz_new / rz_j : 0 ;
// rz_j <- rz_new (active columns) ; p_j = z_j + beta_j p_j (active columns; frozen columns keep p).
// Then the stopping test of the column: it stops when r.M^-1 r <= max(tol2 b.M^-1 b, 1e-30 |b_full|^2 /
// sig) (a NaN residual stops it too; the host reports it).
__global__ void cg_step_p_kernel(double* __restrict__ p, const double* __restrict__ z,
                                 const double* __restrict__ res, double* __restrict__ rz,
                                 double* __restrict__ active, const double* __restrict__ bz,
                                 const double* __restrict__ bfull, double tol2, double floor, int rows) {
  __shared__ double red[NT / 64];
  const int j = blockIdx.x;
  const size_t o = (size_t)rows * j;
  const double rznew = block_coldot(res + o, z + o, rows, red);
  if (active[j] == 0.0) return;
  const double rzold = rz[j];
  const bool act = rzold != 0.0;
  const double beta = act ? rznew / rzold : 0.0;
  for (int i = threadIdx.x; i < rows; i += NT) p[o + i] = z[o + i] + beta * p[o + i];
  __syncthreads();   // every thread has read active[j] before it may be cleared
  if (threadIdx.x == 0) {
    const double r_now = act ? rznew : rzold;
    if (act) rz[j] = rznew;
    const double thr = fmax(fmax(tol2 * fabs(bz[j]), 1e-30 * fabs(bfull[j])), floor);
    if (!(fabs(r_now) > thr)) active[j] = 0.0;
  }
}

}  // namespace

void k_square(const double* in, double* out, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(square_kernel, dim3((n + NT - 1) / NT), dim3(NT), 0, s, in, out, n);
  TNAD_HIP_TRY(hipGetLastError());
}

int resolvent_pcg(Ctx* c, int rows, int r, int k, const double* W, const double* sig, const double* tz,
                  const std::function<void(const double*, double*)>& applyG, const double* b, double* x,
                  double tol, int maxit, double eta, double* sweep_ref, double abs_tol) {
  hipStream_t s = c->stream;
  if (rows < 1 || r < 1 || k < r || k > rows) throw Error(TNAD_ERR_ARG, "resolvent_pcg: bad sizes");
  const size_t nb_ = (size_t)rows * r;
  DevBuf bb(c, nb_ * 8), res(c, nb_ * 8), z(c, nb_ * 8), p(c, nb_ * 8), q(c, nb_ * 8), gv(c, nb_ * 8);
  DevBuf P(c, (size_t)k * r * 8), Cf(c, (size_t)k * r * 8), isig(c, (size_t)r * 8);
  DevBuf sc(c, (size_t)8 * r * 8);   // rz, -, bfull, -, -, -, active, bz: r each (host reads the block)
  double* rz = sc.as<double>();
  double* bfull = rz + 2 * (size_t)r;
  double* active = rz + 6 * (size_t)r;
  double* bz = rz + 7 * (size_t)r;
  const int tr = (r + NT - 1) / NT;

  // V -= Wh (Wh^T V), with the head rows of P as the scratch product
  auto project = [&](double* V) {
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', r, r, rows, rows, rows, k), W, V, P.as<double>(), s);
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', rows, r, r, rows, k, rows, -1.0, 1.0), W, P.as<double>(), V, s);
  };
  // out = M^-1 R
  auto precond = [&](const double* R, double* out) {
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', k, r, rows, rows, rows, k), W, R, P.as<double>(), s);
    hipLaunchKernelGGL(precond_coef_div_kernel, dim3(nb((size_t)k * r + nb_)), dim3(NT), 0, s, P.as<double>(), sig,
                       isig.as<double>(), tz, r, k, eta, Cf.as<double>(), R, rows, out);
    TNAD_HIP_TRY(hipGetLastError());
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', rows, r, k, rows, k, rows, 1.0, 1.0), W, Cf.as<double>(), out, s);
  };
  // out = P (V sig - G V)
  auto op = [&](const double* V, double* out) {
    applyG(V, gv.as<double>());
    hipLaunchKernelGGL(shift_sub_kernel, dim3(nb(nb_)), dim3(NT), 0, s, V, gv.as<double>(), sig, rows, r, out);
    TNAD_HIP_TRY(hipGetLastError());
    project(out);
  };
  auto coldot = [&](const double* X, const double* Y, double* out) {
    hipLaunchKernelGGL(coldot_kernel, dim3(r), dim3(NT), 0, s, X, Y, rows, out);
    TNAD_HIP_TRY(hipGetLastError());
  };

  hipLaunchKernelGGL(inv_sig_kernel, dim3(tr), dim3(NT), 0, s, sig, r, isig.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  k_copy(b, bb.as<double>(), nb_, s);
  coldot(bb.as<double>(), bb.as<double>(), bfull);   // |b_j|^2 / sig_j: the scale of the whole cotangent
  hipLaunchKernelGGL(mul_kernel, dim3(tr), dim3(NT), 0, s, bfull, isig.as<double>(), r);
  project(bb.as<double>());
  precond(bb.as<double>(), x);                          // x0 = M^-1 P b
  coldot(bb.as<double>(), x, bz);                       // b.M^-1 b (stopping scale)
  op(x, q.as<double>());
  k_copy(bb.as<double>(), res.as<double>(), nb_, s);
  k_axpby(-1.0, q.as<double>(), 1.0, res.as<double>(), nb_, s);   // r0 = P b - A x0
  precond(res.as<double>(), z.as<double>());
  k_copy(z.as<double>(), p.as<double>(), nb_, s);
  coldot(res.as<double>(), z.as<double>(), rz);

  // Columns start active; a column stops when r.M^-1 r <= max(tol^2 b.M^-1 b, 1e-30 |b_full|^2 / sig)
  // (the second term measures against the whole right-hand side, so a complement part at the rounding
  // level of b ends at once instead of iterating on noise) or when its curvature p.Ap stops being
  // positive (rounding level reached).  The host applies the test to the starting residual and
  // uploads the active flags once; after that the device applies it at the end of every iteration
  // (the same formula on the same numbers, so the host's copy agrees), and the host only reads the
  // scalar block to stop and to report a NaN.
  std::vector<double> hs((size_t)8 * r), h_act(r, 1.0);
  const double tol2 = tol * tol;
  double floor = 0.0;   // absolute stop level against the sweep's reference
  int it = 0;
  for (;; ++it) {
    c->read_async(hs.data(), sc.ptr, (size_t)8 * r * 8);
    c->sync();
    if (it == 0 && sweep_ref) {
      double mx = 0.0;
      for (int j = 0; j < r; ++j) mx = std::max(mx, std::fabs(hs[(size_t)7 * r + j]));
      if (mx == mx) *sweep_ref = std::max(*sweep_ref, mx);
      if (abs_tol > 0.0) floor = abs_tol * abs_tol * *sweep_ref;
    }
    int nact = 0;
    for (int j = 0; j < r; ++j) {
      const double rzj = hs[j], bzj = hs[(size_t)7 * r + j], bfj = hs[(size_t)2 * r + j];
      if (!(rzj == rzj)) throw Error(TNAD_ERR_NAN, "resolvent_pcg: NaN residual");
      if (it > 0 && hs[(size_t)6 * r + j] == 0.0) h_act[j] = 0.0;
      const double thr = std::max(std::max(tol2 * std::fabs(bzj), 1e-30 * std::fabs(bfj)), floor);
      if (h_act[j] != 0.0 && !(std::fabs(rzj) > thr)) h_act[j] = 0.0;
      nact += h_act[j] != 0.0;
    }
    if (nact == 0 || it >= maxit) break;
    if (it == 0) c->write_async(active, h_act.data(), (size_t)r * 8);
    op(p.as<double>(), q.as<double>());
    hipLaunchKernelGGL(cg_step_xr_kernel, dim3(r), dim3(NT), 0, s, x, res.as<double>(), p.as<double>(),
                       q.as<double>(), rz, active, rows);
    TNAD_HIP_TRY(hipGetLastError());
    precond(res.as<double>(), z.as<double>());
    hipLaunchKernelGGL(cg_step_p_kernel, dim3(r), dim3(NT), 0, s, p.as<double>(), z.as<double>(), res.as<double>(),
                       rz, active, bz, bfull, tol2, floor, rows);
    TNAD_HIP_TRY(hipGetLastError());
  }
  return it;
}

}  // namespace tnad
