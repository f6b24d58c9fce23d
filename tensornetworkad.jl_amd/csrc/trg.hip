// TRG coarse-graining (src/trg.jl:13-44) and its reverse sweep on the device.
//
// Per iteration: max-normalisation (lnZ bookkeeping of src/trg.jl:16-18), the
// two permutations, two truncated SVD splits (trg_svd, src/trg.jl:33-44) and
// the four-tensor contraction (src/trg.jl:25), evaluated pairwise.
//
// SVD: the singular triplets of M (m x m') are read off the symmetric
// augmented matrix B = [[0, M], [M^T, 0]] whose eigenpairs are
// (+-s_i, [u_i; +-v_i] / sqrt(2)); the singular values are therefore not
// squared and keep their relative accuracy.  B is decomposed by the same FP64
// eigensolver as the CTMRG corner matrix (full spectrum up to 192, top-k
// subspace iteration beyond).
//
// Reverse: svd_back (src/trg.jl:72-106) in its SVD form: the pairs among the
// kept triplets by the reference's formula, the rest of the spectrum (and the
// null spaces) by preconditioned resolvent solves with M M^T and M^T M.
#include "trg.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "ctmrg.h"
#include "resolvent.h"

namespace tnad {
namespace {

constexpr int NT = 256;

inline unsigned nb(size_t n) {
  size_t b = (n + NT - 1) / NT;
  if (b > 65535ull * 8) b = 65535ull * 8;
  return (unsigned)(b ? b : 1);
}

// out[j0,j1,j2,j3] = in[i0,i1,i2,i3] where j_k = i_{perm[k]} (Julia permutedims)
__global__ void permute4_kernel(const double* __restrict__ in, int d0, int d1, int d2, int d3, int p0, int p1,
                                int p2, int p3, double* __restrict__ out) {
  const int din[4] = {d0, d1, d2, d3};
  const int p[4] = {p0, p1, p2, p3};
  const int dout[4] = {din[p0], din[p1], din[p2], din[p3]};
  size_t n = (size_t)d0 * d1 * d2 * d3;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int j[4];
    size_t r = e;
    for (int k = 0; k < 4; ++k) { j[k] = r % dout[k]; r /= dout[k]; }
    int i[4];
    for (int k = 0; k < 4; ++k) i[p[k]] = j[k];
    out[e] = in[i[0] + (size_t)d0 * (i[1] + (size_t)d1 * (i[2] + (size_t)d2 * i[3]))];
  }
}

// adjoint of permute4: out[i] += in[j]
__global__ void permute4_back_kernel(const double* __restrict__ gin, int d0, int d1, int d2, int d3, int p0,
                                     int p1, int p2, int p3, double* __restrict__ gout) {
  const int din[4] = {d0, d1, d2, d3};
  const int p[4] = {p0, p1, p2, p3};
  const int dout[4] = {din[p0], din[p1], din[p2], din[p3]};
  size_t n = (size_t)d0 * d1 * d2 * d3;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int j[4];
    size_t r = e;
    for (int k = 0; k < 4; ++k) { j[k] = r % dout[k]; r /= dout[k]; }
    int i[4];
    for (int k = 0; k < 4; ++k) i[p[k]] = j[k];
    gout[i[0] + (size_t)d0 * (i[1] + (size_t)d1 * (i[2] + (size_t)d2 * i[3]))] += gin[e];
  }
}

// one block: max |x| and its first index
__global__ void maxabs_kernel(const double* __restrict__ x, size_t n, double* out_val, int* out_idx) {
  __shared__ double bv[16];
  __shared__ int bi[16];
  double best = -1.0;
  int idx = 0x7fffffff;
  for (size_t e = threadIdx.x; e < n; e += blockDim.x) {
    double v = fabs(x[e]);
    if (v > best) { best = v; idx = (int)e; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    double ob = __shfl_down(best, o);
    int oi = __shfl_down(idx, o);
    if (ob > best || (ob == best && oi < idx)) { best = ob; idx = oi; }
  }
  if ((threadIdx.x & 63) == 0) { bv[threadIdx.x >> 6] = best; bi[threadIdx.x >> 6] = idx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = bv[0];
    int ii = bi[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      if (bv[w] > b || (bv[w] == b && bi[w] < ii)) { b = bv[w]; ii = bi[w]; }
    *out_val = b;
    *out_idx = ii;
  }
}

// B = [[0, M], [M^T, 0]]  (N = m + mp)
__global__ void augment_kernel(const double* __restrict__ M, int m, int mp, double* __restrict__ B) {
  const int N = m + mp;
  size_t tot = (size_t)N * N;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    int r = e % N, c = e / N;
    double v = 0.0;
    if (r < m && c >= m) v = M[r + (size_t)m * (c - m)];
    else if (r >= m && c < m) v = M[c + (size_t)m * (r - m)];
    B[e] = v;
  }
}

// U[:, j] = sqrt(2) X[0:m, col_j], V[:, j] = sqrt(2) X[m:, col_j]  (col_j < 0: zero column)
__global__ void split_vectors_kernel(const double* __restrict__ X, int N, int m, const int* __restrict__ cols,
                                     int kk, double* __restrict__ U, double* __restrict__ V) {
  const int mp = N - m;
  size_t tot = (size_t)N * kk;
  const double r2 = 1.4142135623730951;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    int r = e % N, j = e / N;
    int cidx = cols[j];
    double v = cidx >= 0 ? r2 * X[r + (size_t)N * cidx] : 0.0;
    if (r < m) U[r + (size_t)m * j] = v;
    else V[(r - m) + (size_t)mp * j] = v;
  }
}

// u_out[:, j] = U[:, j] sqrt(s_j) (m x k);  v_out[j, :] = sqrt(s_j) V[:, j]^T (rows j < k of an
// ldv x mp column-major array)
__global__ void scale_split_kernel(const double* __restrict__ U, const double* __restrict__ V,
                                   const double* __restrict__ sq, int m, int mp, int k, double* __restrict__ uo,
                                   double* __restrict__ vo, int ldv) {
  size_t tot = (size_t)(m + mp) * k;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    if (e < (size_t)m * k) {
      int j = e / m;
      uo[e] = U[e] * sq[j];
    } else {
      size_t f = e - (size_t)m * k;
      int c = f % mp, j = f / mp;
      vo[j + (size_t)ldv * c] = sq[j] * V[c + (size_t)mp * j];
    }
  }
}

__global__ void trace_kernel(const double* __restrict__ a, int d0, int d1, double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int e = threadIdx.x; e < d0 * d1; e += blockDim.x) {
    int i = e % d0, j = e / d0;
    acc += a[i + (size_t)d0 * (j + (size_t)d1 * (i + (size_t)d0 * j))];
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    *out = s;
  }
}

__global__ void trace_back_kernel(double* __restrict__ abar, int d0, int d1, double val) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d0 * d1; e += gridDim.x * blockDim.x) {
    int i = e % d0, j = e / d0;
    abar[i + (size_t)d0 * (j + (size_t)d1 * (i + (size_t)d0 * j))] = val;
  }
}

// dU[:, j] = W[0:m, j] / sqrt(2), dV[:, j] = W[m:, j] / sqrt(2)
__global__ void unsplit_kernel(const double* __restrict__ W, int m, int mp, int kk, double* __restrict__ dU,
                               double* __restrict__ dV) {
  const int N = m + mp;
  const double ir2 = 0.7071067811865476;
  size_t tot = (size_t)N * kk;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    int r = e % N, j = e / N;
    if (r < m) dU[r + (size_t)m * j] = W[e] * ir2;
    else dV[(r - m) + (size_t)mp * j] = W[e] * ir2;
  }
}

// X[:, j] *= s[j]
__global__ void scale_cols_by_kernel(double* __restrict__ X, int rows, int cols, const double* __restrict__ sv) {
  size_t tot = (size_t)rows * cols;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x)
    X[e] *= sv[e / rows];
}

// R (K x K) of svd_back (src/trg.jl:72-106) restricted to the K computed singular triplets,
// cotangents dU, dV on the first kk columns (gu = U^T dU, gv = V^T dV, K x kk), dS on kk:
//   i,j < kk : (F_ij gu_ij + F_ji gu_ji) s_j + s_i (F_ij gv_ij + F_ji gv_ji) + delta_ij dS_j
//   i >= kk, j < kk : F_ij (gu_ij s_j + s_i gv_ij)
//   i < kk, j >= kk : F_ji (gu_ji s_j + s_i gv_ji)
// F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eta), formed from the rounded squares (no FP
// contraction) so that equal singular values give exactly F = 0 as in the reference.
__global__ void svdback_R_kernel(const double* __restrict__ gu, const double* __restrict__ gv,
                                 const double* __restrict__ sv, const double* __restrict__ dS, int K, int kk,
                                 double eta, double* __restrict__ R) {
#pragma clang fp contract(off)
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= K * K) return;
  int i = e % K, j = e / K;
  auto F = [&](int a, int b) {   // F_ab = (s_b^2 - s_a^2) / ((s_b^2 - s_a^2)^2 + eta)
    double sa2 = sv[a] * sv[a];
    double sb2 = sv[b] * sv[b];
    double d = sb2 - sa2;
    double d2 = d * d;
    return d / (d2 + eta);
  };
  double v = 0.0;
  if (i < kk && j < kk) {
    double gu_ij = gu[i + (size_t)K * j], gu_ji = gu[j + (size_t)K * i];
    double gv_ij = gv[i + (size_t)K * j], gv_ji = gv[j + (size_t)K * i];
    double fij = F(i, j), fji = F(j, i);
    double t1 = fij * gu_ij + fji * gu_ji;
    double t2 = fij * gv_ij + fji * gv_ji;
    v = t1 * sv[j] + sv[i] * t2;
    if (i == j) v += dS[j];
  } else if (j < kk) {
    double t = gu[i + (size_t)K * j] * sv[j] + sv[i] * gv[i + (size_t)K * j];
    v = F(i, j) * t;
  } else if (i < kk) {
    double t = gu[j + (size_t)K * i] * sv[j] + sv[i] * gv[j + (size_t)K * i];
    v = F(j, i) * t;
  }
  R[e] = v;
}


// gradient bookkeeping of u = U sqrt(s), v = sqrt(s) V^T:
// Wbar[0:m, j] = sqrt(2) ubar[:, j] sqrt(s_j), Wbar[m:, j] = sqrt(2) vbar[j, :]^T sqrt(s_j)
// sbar_j = (ubar[:, j] . U[:, j] + vbar[j, :] . V[:, j]) / (2 sqrt(s_j))  (0 for s_j = 0)
__global__ void svd_cotangent_kernel(const double* __restrict__ ubar, const double* __restrict__ vbar,
                                     const double* __restrict__ U, const double* __restrict__ V,
                                     const double* __restrict__ sq, int m, int mp, int k,
                                     double* __restrict__ Wbar, double* __restrict__ sbar) {
  const int j = blockIdx.x;
  const int N = m + mp;
  const double r2 = 1.4142135623730951;
  __shared__ double red[4];
  double acc = 0.0;
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    if (r < m) {
      double ub = ubar[r + (size_t)m * j];
      Wbar[r + (size_t)N * j] = r2 * ub * sq[j];
      acc += ub * U[r + (size_t)m * j];
    } else {
      int c = r - m;
      double vb = vbar[j + (size_t)k * c];
      Wbar[r + (size_t)N * j] = r2 * vb * sq[j];
      acc += vb * V[c + (size_t)mp * j];
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0] + red[1] + red[2] + red[3];
    sbar[j] = sq[j] > 0.0 ? t / (2.0 * sq[j]) : 0.0;
  }
}


}  // namespace

// ---------------------------------------------------------------------------
// truncated SVD split of one rank-4 tensor
// ---------------------------------------------------------------------------
struct SvdTape {
  int d1 = 0, d2 = 0, d3 = 0, d4 = 0, m = 0, mp = 0, N = 0, k = 0, K = 0;
  DevBuf M;                    // the m x mp matrix being split
  DevBuf U, V, sv, sq;         // the K computed singular triplets (U m x K, V mp x K), s and sqrt(s)
  std::vector<double> hs;
};

// singular triplets of M from the top eigenpairs of [[0, M], [M^T, 0]]
static void svd_split(Ctx* c, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol, double* u_out,
                      double* v_out, SvdTape& tp) {
  hipStream_t s = c->stream;
  const int m = d1 * d2, mp = d3 * d4, N = m + mp;
  const int len = std::min(m, mp);
  tp.d1 = d1; tp.d2 = d2; tp.d3 = d3; tp.d4 = d4; tp.m = m; tp.mp = mp; tp.N = N;
  tp.M = DevBuf(c, (size_t)m * mp * 8);
  k_copy(t, tp.M.as<double>(), (size_t)m * mp, s);
  DevBuf B(c, (size_t)N * N * 8);
  hipLaunchKernelGGL(augment_kernel, dim3(nb((size_t)N * N)), dim3(NT), 0, s, t, m, mp, B.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  const int want = std::min(N, 2 * std::min(dmax, len) + 2);
  int keig = (N <= JACOBI_MAX) ? N : std::min(JACOBI_MAX, std::max(want + 24, 2 * want));
  keig = std::min(keig, N);
  DevBuf X(c, (size_t)N * keig * 8), AX(c, (size_t)N * keig * 8), theta(c, (size_t)keig * 8);
  keig = topk_eig(c, B.as<double>(), N, std::min(want, keig), keig, nullptr, 300, 1e-13, X.as<double>(),
                  AX.as<double>(), theta.as<double>(), nullptr);
  std::vector<double> th(keig);
  c->read_async(th.data(), theta.ptr, (size_t)keig * 8);
  c->sync();
  // singular values = the positive eigenvalues in descending order; values at the eigensolver's
  // rounding level (relative 1e-14) are not resolved as +-s pairs and count as exact zeros
  const double zero_level = 1e-14 * std::fabs(th[0]);
  std::vector<int> pos;
  tp.hs.clear();
  for (int i = 0; i < keig && (int)pos.size() < len; ++i)
    if (th[i] > zero_level) { pos.push_back(i); tp.hs.push_back(th[i]); }
  const int K = (int)pos.size();
  int above = 0;
  for (int i = 0; i < K; ++i) above += tp.hs[i] > tol;
  // keep = min(#{s > tol} + 1, dmax, length(s))   (src/trg.jl:37); kept values beyond the
  // resolved positive spectrum are zero singular values and give zero columns
  const int k = std::min(std::min(above + 1, dmax), len);
  tp.k = k;
  tp.K = K;
  const int Kc = std::max(K, 1);
  std::vector<double> hsq(Kc, 0.0);
  for (int j = 0; j < K; ++j) hsq[j] = std::sqrt(tp.hs[j]);
  DevBuf cols(c, (size_t)Kc * sizeof(int) + 16);
  tp.sv = DevBuf(c, (size_t)Kc * 8);
  tp.sq = DevBuf(c, (size_t)std::max(k, Kc) * 8);
  tp.U = DevBuf(c, (size_t)m * Kc * 8);
  tp.V = DevBuf(c, (size_t)mp * Kc * 8);
  k_fill(tp.sq.as<double>(), (size_t)std::max(k, Kc), 0.0, s);
  if (K > 0) {
    c->write_async(cols.ptr, pos.data(), (size_t)K * sizeof(int));
    c->write_async(tp.sv.ptr, tp.hs.data(), (size_t)K * 8);
    c->write_async(tp.sq.ptr, hsq.data(), (size_t)K * 8);
    hipLaunchKernelGGL(split_vectors_kernel, dim3(nb((size_t)N * K)), dim3(NT), 0, s, X.as<double>(), N, m,
                       cols.as<int>(), K, tp.U.as<double>(), tp.V.as<double>());
  }
  // u = U[:, 1:k] sqrt(s), v = sqrt(s) V[:, 1:k]^T; columns j >= K are zero
  k_fill(u_out, (size_t)m * k, 0.0, s);
  k_fill(v_out, (size_t)k * mp, 0.0, s);
  const int kv = std::min(k, K);
  if (kv > 0) {
    hipLaunchKernelGGL(scale_split_kernel, dim3(nb((size_t)(m + mp) * kv)), dim3(NT), 0, s, tp.U.as<double>(),
                       tp.V.as<double>(), tp.sq.as<double>(), m, mp, kv, u_out, v_out, k);
    TNAD_HIP_TRY(hipGetLastError());
  }
  c->sync();
}

// svd_back (src/trg.jl:72-106) for u = U sqrt(s), v = sqrt(s) V^T: cotangents of u (m x k) and
// v (k x mp) -> cotangent of M (m x mp).  Pairs among the kk kept, resolved triplets use the
// reference's formula verbatim (R, kk x kk).  Every other singular pair, and the null spaces
// of a rectangular or rank-deficient M, enter through
//   x_j = P_U (s_j^2 - M M^T)^-1 P_U (s_j dU_j + M dV_j)   (times v_j^T)
//   y_j = P_V (s_j^2 - M^T M)^-1 P_V (M^T dU_j + s_j dV_j)  (u_j times y_j^T)
// with P_U, P_V the projectors off the kept vectors, solved by resolvent_pcg with the
// further computed triplets as preconditioner.  Kept columns beyond the resolved spectrum
// (zero singular values, zero columns of u and v) carry no cotangent.
static void svd_split_back(Ctx* c, const SvdTape& tp, const double* ubar, const double* vbar, double* tbar) {
  hipStream_t s = c->stream;
  const int m = tp.m, mp = tp.mp, k = tp.k, K = tp.K;
  const int kk = std::min(k, K);            // kept columns carrying a resolved singular triplet
  k_fill(tbar, (size_t)m * mp, 0.0, s);
  if (kk == 0) return;
  const double* U = tp.U.as<double>();
  const double* V = tp.V.as<double>();
  // dU = ubar sqrt(s), dV = vbar^T sqrt(s), dS = (ubar . U + vbar . V) / (2 sqrt(s))
  DevBuf dU(c, (size_t)m * kk * 8), dV(c, (size_t)mp * kk * 8), dS(c, (size_t)kk * 8), Wtmp(c, (size_t)(m + mp) * kk * 8);
  hipLaunchKernelGGL(svd_cotangent_kernel, dim3(kk), dim3(NT), 0, s, ubar, vbar, U, V, tp.sq.as<double>(), m, mp, k,
                     Wtmp.as<double>(), dS.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  // svd_cotangent_kernel writes sqrt(2) * [dU; dV] per column (ld m + mp): undo the sqrt(2) while splitting
  const int N = m + mp;
  hipLaunchKernelGGL(unsplit_kernel, dim3(nb((size_t)N * kk)), dim3(NT), 0, s, Wtmp.as<double>(), m, mp, kk,
                     dU.as<double>(), dV.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  DevBuf gu(c, (size_t)kk * kk * 8), gv(c, (size_t)kk * kk * 8), R(c, (size_t)kk * kk * 8), UR(c, (size_t)m * kk * 8);
  tc_launch(c, Prec::F64, gemm_desc('T', 'N', kk, kk, m, m, m, kk), U, dU.as<double>(), gu.as<double>(), s);
  tc_launch(c, Prec::F64, gemm_desc('T', 'N', kk, kk, mp, mp, mp, kk), V, dV.as<double>(), gv.as<double>(), s);
  hipLaunchKernelGGL(svdback_R_kernel, dim3((kk * kk + NT - 1) / NT), dim3(NT), 0, s, gu.as<double>(),
                     gv.as<double>(), tp.sv.as<double>(), dS.as<double>(), kk, kk, 1e-40, R.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  // Mbar = U_kk R V_kk^T
  tc_launch(c, Prec::F64, gemm_desc('N', 'N', m, kk, kk, m, kk, m), U, R.as<double>(), UR.as<double>(), s);
  tc_launch(c, Prec::F64, gemm_desc('N', 'T', m, mp, kk, m, mp, m), UR.as<double>(), V, tbar, s);
  if (m == kk && mp == kk) return;   // square, fully kept: nothing outside
  const double* M = tp.M.as<double>();
  DevBuf sq2(c, (size_t)K * 8);
  k_square(tp.sv.as<double>(), sq2.as<double>(), K, s);
  const double* sig = sq2.as<double>();
  const double* tz = sq2.as<double>() + kk;
  DevBuf T1(c, (size_t)std::max(m, mp) * kk * 8);
  if (m > kk) {
    // left: b = s dU + M dV ; G = M M^T
    DevBuf bl(c, (size_t)m * kk * 8), X(c, (size_t)m * kk * 8);
    k_copy(dU.as<double>(), bl.as<double>(), (size_t)m * kk, s);
    hipLaunchKernelGGL(scale_cols_by_kernel, dim3(nb((size_t)m * kk)), dim3(NT), 0, s, bl.as<double>(), m, kk,
                       tp.sv.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
    tc_launch(c, Prec::F64, gemm_desc('N', 'N', m, kk, mp, m, mp, m, 1.0, 1.0), M, dV.as<double>(), bl.as<double>(), s);
    auto applyMMt = [&](const double* Vb, double* out) {
      tc_launch(c, Prec::F64, gemm_desc('T', 'N', mp, kk, m, m, m, mp), M, Vb, T1.as<double>(), s);
      tc_launch(c, Prec::F64, gemm_desc('N', 'N', m, kk, mp, m, mp, m), M, T1.as<double>(), out, s);
    };
    resolvent_pcg(c, m, kk, K, U, sig, tz, applyMMt, bl.as<double>(), X.as<double>(), 1e-13, 200);
    // Mbar += X V_kk^T
    tc_launch(c, Prec::F64, gemm_desc('N', 'T', m, mp, kk, m, mp, m, 1.0, 1.0), X.as<double>(), V, tbar, s);
  }
  if (mp > kk) {
    // right: b' = M^T dU + s dV ; G = M^T M
    DevBuf br(c, (size_t)mp * kk * 8), Y(c, (size_t)mp * kk * 8);
    k_copy(dV.as<double>(), br.as<double>(), (size_t)mp * kk, s);
    hipLaunchKernelGGL(scale_cols_by_kernel, dim3(nb((size_t)mp * kk)), dim3(NT), 0, s, br.as<double>(), mp, kk,
                       tp.sv.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
    tc_launch(c, Prec::F64, gemm_desc('T', 'N', mp, kk, m, m, m, mp, 1.0, 1.0), M, dU.as<double>(), br.as<double>(), s);
    auto applyMtM = [&](const double* Vb, double* out) {
      tc_launch(c, Prec::F64, gemm_desc('N', 'N', m, kk, mp, m, mp, m), M, Vb, T1.as<double>(), s);
      tc_launch(c, Prec::F64, gemm_desc('T', 'N', mp, kk, m, m, m, mp), M, T1.as<double>(), out, s);
    };
    resolvent_pcg(c, mp, kk, K, V, sig, tz, applyMtM, br.as<double>(), Y.as<double>(), 1e-13, 200);
    // Mbar += U_kk Y^T
    tc_launch(c, Prec::F64, gemm_desc('N', 'T', m, mp, kk, m, mp, m, 1.0, 1.0), U, Y.as<double>(), tbar, s);
  }
}
// ---------------------------------------------------------------------------
// trg (src/trg.jl:13-30)
// ---------------------------------------------------------------------------
struct TrgIter {
  int du = 0, dr = 0;              // dims of the input (u=d, r=l)
  DevBuf a_in;                     // tensor before normalisation
  double maxval = 1.0;
  int idx = 0;
  double sgn = 1.0;
  SvdTape s1, s2;
  DevBuf dr_t, ul_t, ld_t, ru_t;   // the four factors
  int k1 = 0, k2 = 0;
};

void trg_run(Ctx* c, const double* a_host, int q, int chi, int niter, double tol, double* lnz_out, double* abar_out) {
  hipStream_t s = c->stream;
  if (q < 1 || chi < 1 || niter < 0) throw Error(TNAD_ERR_ARG, "trg: bad sizes");
  if ((long long)chi * chi > 4096) throw Error(TNAD_ERR_UNSUPPORTED, "trg: chi above 64 is not supported");
  const bool grad = abar_out != nullptr;
  int du = q, dr = q;   // a is (du, dr, du, dr)
  DevBuf a(c, (size_t)q * q * q * q * 8);
  copy_in(c, a.ptr, a_host, (size_t)q * q * q * q * 8);
  std::vector<std::unique_ptr<TrgIter>> tape;
  double lnz = 0.0;
  DevBuf mv(c, 16), mi(c, 16);
  for (int it = 1; it <= niter; ++it) {
    const size_t na = (size_t)du * dr * du * dr;
    std::unique_ptr<TrgIter> T(new TrgIter());
    T->du = du; T->dr = dr;
    // maxval = maximum(abs.(a)); a /= maxval; lnZ += 2^(1-n) log(maxval)     (src/trg.jl:16-18)
    hipLaunchKernelGGL(maxabs_kernel, dim3(1), dim3(1024), 0, s, a.as<double>(), na, mv.as<double>(), mi.as<int>());
    TNAD_HIP_TRY(hipGetLastError());
    double hv = 0.0;
    int hi = 0;
    c->read_async(&hv, mv.ptr, 8);
    c->read_async(&hi, mi.ptr, sizeof(int));
    c->sync();
    if (!(hv > 0.0)) throw Error(TNAD_ERR_NAN, "trg: tensor vanished or is not finite");
    if (grad) {
      T->a_in = DevBuf(c, na * 8);
      k_copy(a.as<double>(), T->a_in.as<double>(), na, s);
      double v = 0.0;
      c->read_async(&v, a.as<double>() + hi, 8);
      c->sync();
      T->sgn = v < 0 ? -1.0 : 1.0;
    }
    T->maxval = hv;
    T->idx = hi;
    k_axpby(1.0 / hv, a.as<double>(), 0.0, a.as<double>(), na, s);
    lnz += std::pow(2.0, 1 - it) * std::log(hv);
    // dr_ul = ein"urdl -> drul"(a) ; ld_ru = ein"urdl -> ldru"(a)       (src/trg.jl:20-21)
    DevBuf p1(c, na * 8), p2(c, na * 8);
    hipLaunchKernelGGL(permute4_kernel, dim3(nb(na)), dim3(NT), 0, s, a.as<double>(), du, dr, du, dr, 2, 1, 0, 3,
                       p1.as<double>());
    hipLaunchKernelGGL(permute4_kernel, dim3(nb(na)), dim3(NT), 0, s, a.as<double>(), du, dr, du, dr, 3, 2, 1, 0,
                       p2.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
    // dr, ul = trg_svd(dr_ul, chi, tol) ; ld, ru = trg_svd(ld_ru, chi, tol)  (src/trg.jl:22-23)
    const int kmax1 = std::min(chi, du * dr), kmax2 = std::min(chi, dr * du);
    T->dr_t = DevBuf(c, (size_t)du * dr * kmax1 * 8 + 8);
    T->ul_t = DevBuf(c, (size_t)kmax1 * du * dr * 8 + 8);
    T->ld_t = DevBuf(c, (size_t)dr * du * kmax2 * 8 + 8);
    T->ru_t = DevBuf(c, (size_t)kmax2 * dr * du * 8 + 8);
    svd_split(c, p1.as<double>(), du, dr, du, dr, chi, tol, T->dr_t.as<double>(), T->ul_t.as<double>(), T->s1);
    svd_split(c, p2.as<double>(), dr, du, dr, du, chi, tol, T->ld_t.as<double>(), T->ru_t.as<double>(), T->s2);
    const int k1 = T->s1.k, k2 = T->s2.k;
    T->k1 = k1; T->k2 = k2;
    // a = ein"npu,por,dom,lmn -> urdl"(dr, ld, ul, ru)                   (src/trg.jl:25)
    const long long n_ = du, p_ = dr, o_ = du, m_ = dr, K1 = k1, K2 = k2;
    DevBuf Xb(c, (size_t)n_ * K1 * o_ * K2 * 8), Yb(c, (size_t)K1 * o_ * K2 * n_ * 8),
        an(c, (size_t)K1 * K2 * K1 * K2 * 8);
    tc(c, Prec::F64, cm("npu", {n_, p_, K1}), T->dr_t.ptr, cm("por", {p_, o_, K2}), T->ld_t.ptr,
       cm("nuor", {n_, K1, o_, K2}), Xb.ptr);
    tc(c, Prec::F64, cm("dom", {K1, o_, m_}), T->ul_t.ptr, cm("lmn", {K2, m_, n_}), T->ru_t.ptr,
       cm("doln", {K1, o_, K2, n_}), Yb.ptr);
    tc(c, Prec::F64, cm("nuor", {n_, K1, o_, K2}), Xb.ptr, cm("doln", {K1, o_, K2, n_}), Yb.ptr,
       cm("urdl", {K1, K2, K1, K2}), an.ptr);
    a = std::move(an);
    du = k1;
    dr = k2;
    if (grad) tape.push_back(std::move(T));
  }
  // trace = ein"ijij -> "(a)[] ; lnZ += log(trace) / 2^niter               (src/trg.jl:27-28)
  DevBuf tr(c, 16);
  hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(256), 0, s, a.as<double>(), du, dr, tr.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  double htr = 0.0;
  c->read_async(&htr, tr.ptr, 8);
  c->sync();
  if (!(htr > 0.0)) throw Error(TNAD_ERR_NAN, "trg: non-positive trace");
  lnz += std::log(htr) / std::pow(2.0, niter);
  if (lnz_out) *lnz_out = lnz;
  if (!grad) return;

  // ---- reverse sweep -------------------------------------------------------
  DevBuf abar(c, (size_t)du * dr * du * dr * 8);
  k_fill(abar.as<double>(), (size_t)du * dr * du * dr, 0.0, s);
  hipLaunchKernelGGL(trace_back_kernel, dim3(nb((size_t)du * dr)), dim3(NT), 0, s, abar.as<double>(), du, dr,
                     1.0 / (std::pow(2.0, niter) * htr));
  TNAD_HIP_TRY(hipGetLastError());
  for (int it = niter; it >= 1; --it) {
    TrgIter& T = *tape[it - 1];
    const int du0 = T.du, dr0 = T.dr, k1 = T.k1, k2 = T.k2;
    const long long n_ = du0, p_ = dr0, o_ = du0, m_ = dr0, K1 = k1, K2 = k2;
    const size_t na = (size_t)du0 * dr0 * du0 * dr0;
    // recompute the pairwise intermediates
    DevBuf Xb(c, (size_t)n_ * K1 * o_ * K2 * 8), Yb(c, (size_t)K1 * o_ * K2 * n_ * 8);
    tc(c, Prec::F64, cm("npu", {n_, p_, K1}), T.dr_t.ptr, cm("por", {p_, o_, K2}), T.ld_t.ptr,
       cm("nuor", {n_, K1, o_, K2}), Xb.ptr);
    tc(c, Prec::F64, cm("dom", {K1, o_, m_}), T.ul_t.ptr, cm("lmn", {K2, m_, n_}), T.ru_t.ptr,
       cm("doln", {K1, o_, K2, n_}), Yb.ptr);
    DevBuf Xbar(c, (size_t)n_ * K1 * o_ * K2 * 8), Ybar(c, (size_t)K1 * o_ * K2 * n_ * 8);
    tc(c, Prec::F64, cm("urdl", {K1, K2, K1, K2}), abar.ptr, cm("doln", {K1, o_, K2, n_}), Yb.ptr,
       cm("nuor", {n_, K1, o_, K2}), Xbar.ptr);
    tc(c, Prec::F64, cm("nuor", {n_, K1, o_, K2}), Xb.ptr, cm("urdl", {K1, K2, K1, K2}), abar.ptr,
       cm("doln", {K1, o_, K2, n_}), Ybar.ptr);
    DevBuf drb(c, (size_t)n_ * p_ * K1 * 8 + 8), ldb(c, (size_t)p_ * o_ * K2 * 8 + 8),
        ulb(c, (size_t)K1 * o_ * m_ * 8 + 8), rub(c, (size_t)K2 * m_ * n_ * 8 + 8);
    tc(c, Prec::F64, cm("nuor", {n_, K1, o_, K2}), Xbar.ptr, cm("por", {p_, o_, K2}), T.ld_t.ptr,
       cm("npu", {n_, p_, K1}), drb.ptr);
    tc(c, Prec::F64, cm("npu", {n_, p_, K1}), T.dr_t.ptr, cm("nuor", {n_, K1, o_, K2}), Xbar.ptr,
       cm("por", {p_, o_, K2}), ldb.ptr);
    tc(c, Prec::F64, cm("doln", {K1, o_, K2, n_}), Ybar.ptr, cm("lmn", {K2, m_, n_}), T.ru_t.ptr,
       cm("dom", {K1, o_, m_}), ulb.ptr);
    tc(c, Prec::F64, cm("dom", {K1, o_, m_}), T.ul_t.ptr, cm("doln", {K1, o_, K2, n_}), Ybar.ptr,
       cm("lmn", {K2, m_, n_}), rub.ptr);
    // through the two SVD splits
    DevBuf t1b(c, na * 8), t2b(c, na * 8);
    svd_split_back(c, T.s1, drb.as<double>(), ulb.as<double>(), t1b.as<double>());
    svd_split_back(c, T.s2, ldb.as<double>(), rub.as<double>(), t2b.as<double>());
    // permutation adjoints into the normalised tensor's cotangent
    DevBuf anb(c, na * 8);
    k_fill(anb.as<double>(), na, 0.0, s);
    hipLaunchKernelGGL(permute4_back_kernel, dim3(nb(na)), dim3(NT), 0, s, t1b.as<double>(), du0, dr0, du0, dr0, 2,
                       1, 0, 3, anb.as<double>());
    hipLaunchKernelGGL(permute4_back_kernel, dim3(nb(na)), dim3(NT), 0, s, t2b.as<double>(), du0, dr0, du0, dr0, 3,
                       2, 1, 0, anb.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
    // a_n = a_in / maxval ; lnZ += 2^(1-n) log(maxval)
    DevBuf red(c, 300 * 8);
    k_dot(anb.as<double>(), T.a_in.as<double>(), na, red.as<double>(), s);
    double dotv = 0.0;
    c->read_async(&dotv, red.ptr, 8);
    c->sync();
    const double mv_ = T.maxval;
    const double mbar = std::pow(2.0, 1 - it) / mv_ - dotv / (mv_ * mv_);
    DevBuf ain(c, na * 8);
    k_axpby(1.0 / mv_, anb.as<double>(), 0.0, ain.as<double>(), na, s);
    // maximum(abs.(a)) sends its cotangent to the first maximiser with the sign of that entry
    std::vector<double> one(1);
    c->read_async(one.data(), ain.as<double>() + T.idx, 8);
    c->sync();
    one[0] += mbar * T.sgn;
    c->write_async(ain.as<double>() + T.idx, one.data(), 8);
    abar = std::move(ain);
  }
  copy_out(c, abar_out, abar.ptr, (size_t)q * q * q * q * 8);
}

void trg_svd_run(Ctx* c, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol, double* u,
                 double* v, int* k_out) {
  if (d1 < 1 || d2 < 1 || d3 < 1 || d4 < 1 || dmax < 1) throw Error(TNAD_ERR_ARG, "trg_svd: bad sizes");
  const int m = d1 * d2, mp = d3 * d4;
  enforce(trg_split_limit(m, mp, dmax));
  const int kmax = std::min(dmax, std::min(m, mp));
  DevBuf bt(c, (size_t)m * mp * 8), bu(c, (size_t)m * kmax * 8 + 8), bv(c, (size_t)kmax * mp * 8 + 8);
  copy_in(c, bt.ptr, t, (size_t)m * mp * 8);
  SvdTape tp;
  svd_split(c, bt.as<double>(), d1, d2, d3, d4, dmax, tol, bu.as<double>(), bv.as<double>(), tp);
  const int k = tp.k;
  // outputs are written with the caller's capacity kmax: u (m x kmax), v (kmax x mp)
  std::vector<double> hu((size_t)m * k), hv((size_t)k * mp);
  c->read_async(hu.data(), bu.ptr, (size_t)m * k * 8);
  c->read_async(hv.data(), bv.ptr, (size_t)k * mp * 8);
  c->sync();
  std::vector<double> ou((size_t)m * kmax, 0.0), ov((size_t)kmax * mp, 0.0);
  for (size_t i = 0; i < (size_t)m * k; ++i) ou[i] = hu[i];
  for (int j = 0; j < k; ++j)
    for (int cc = 0; cc < mp; ++cc) ov[j + (size_t)kmax * cc] = hv[j + (size_t)k * cc];
  copy_out(c, u, ou.data(), ou.size() * 8);
  copy_out(c, v, ov.data(), ov.size() * 8);
  c->sync();   // ou and ov are the copies' sources
  if (k_out) *k_out = k;
}

// Pullback of trg_svd (src/trg.jl:33-44 through the svd adjoint of src/trg.jl:55-61,72-106): the split is
// recomputed from t (the same singular triplets as the forward), then svd_split_back maps the cotangents
// of u (d1,d2,k) and v (k,d3,d4) to that of t.  k must be the forward's kept count.
void trg_svd_back_run(Ctx* c, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol,
                      const double* ubar, const double* vbar, int k, double* tbar) {
  if (d1 < 1 || d2 < 1 || d3 < 1 || d4 < 1 || dmax < 1 || k < 1) throw Error(TNAD_ERR_ARG, "trg_svd_back: bad sizes");
  const int m = d1 * d2, mp = d3 * d4;
  enforce(trg_split_limit(m, mp, dmax));
  const int kmax = std::min(dmax, std::min(m, mp));
  DevBuf bt(c, (size_t)m * mp * 8), bu(c, (size_t)m * kmax * 8 + 8), bv(c, (size_t)kmax * mp * 8 + 8);
  copy_in(c, bt.ptr, t, (size_t)m * mp * 8);
  SvdTape tp;
  svd_split(c, bt.as<double>(), d1, d2, d3, d4, dmax, tol, bu.as<double>(), bv.as<double>(), tp);
  if (tp.k != k) throw Error(TNAD_ERR_SHAPE, "trg_svd_back: k is not the forward's kept count");
  DevBuf bub(c, (size_t)m * k * 8), bvb(c, (size_t)k * mp * 8), btb(c, (size_t)m * mp * 8);
  copy_in(c, bub.ptr, ubar, (size_t)m * k * 8);
  copy_in(c, bvb.ptr, vbar, (size_t)k * mp * 8);
  svd_split_back(c, tp, bub.as<double>(), bvb.as<double>(), btb.as<double>());
  copy_out(c, tbar, btb.ptr, (size_t)m * mp * 8);
  c->sync();
}

}  // namespace tnad
