// Small dense FP64 kernels for the eigensolver, one workgroup each (m <= 192).
//
//  * jacobi_rot_kernel / jacobi_apply_kernel : cyclic two-sided Jacobi with the
//    round-robin (circle) ordering on a packed triangle in LDS (192x192 FP64 =
//    148 KB of the 160 KB a CU offers), so each round applies m/2 disjoint
//    rotations in parallel and every 2x2 block is updated by one thread; the
//    rotations are recorded and replayed on the identity by many workgroups.
//    Used for the full spectrum when n <= 192 and for the Rayleigh-Ritz problem
//    of the block subspace iteration.
//  * chol_inv_kernel : Cholesky of a (shifted) Gram matrix followed by the
//    triangular inverse, register-resident, for CholeskyQR orthonormalisation.
#include "tnad_internal.h"

namespace tnad {
namespace {

constexpr int EIG_NT = 1024;
// 2x2 blocks per thread per Jacobi round: ceil(P (P + 1) / 2 / EIG_NT) for P = JACOBI_MAX / 2
constexpr int JB_MAX = ((JACOBI_MAX / 2) * (JACOBI_MAX / 2 + 1) / 2 + EIG_NT - 1) / EIG_NT;

__device__ __forceinline__ int pk(int i, int j) {  // packed upper, i <= j, column-major
  return j * (j + 1) / 2 + i;
}
__device__ __forceinline__ int pku(int i, int j) { return i <= j ? pk(i, j) : pk(j, i); }

__device__ __forceinline__ void decode_tri(int b, int& i, int& j) {
  // b = j(j+1)/2 + i with 0 <= i <= j
  int jj = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
  while ((jj + 1) * (jj + 2) / 2 <= b) ++jj;
  while (jj * (jj + 1) / 2 > b) --jj;
  j = jj;
  i = b - jj * (jj + 1) / 2;
}

__device__ __forceinline__ void rr_pair(int r_in, int t, int mm, int& p, int& q) {
  // round-robin (circle) pairing of round r_in, slot t
  if (t == 0) { p = r_in; q = mm - 1; }
  else { p = (r_in + t) % (mm - 1); q = (r_in - t + (mm - 1)) % (mm - 1); }
  if (p > q) { int tmp = p; p = q; q = tmp; }
}

// Jacobi sweeps on the packed matrix in LDS.  The rotations of every round are
// recorded (rot[(round * P + slot) * 2 + {0,1}] = {c, s}); the eigenvector
// matrix is formed afterwards by jacobi_apply_kernel.  Writes the (unsorted)
// diagonal to diag and the number of rounds to *rounds_out.
__global__ __launch_bounds__(EIG_NT) void jacobi_rot_kernel(const double* __restrict__ H, int ldh, int m,
                                                            double* __restrict__ rot, int max_sweeps,
                                                            double* __restrict__ diag, int* rounds_out,
                                                            int* sweeps_out) {
  extern __shared__ double smem[];
  const int mm = m + (m & 1);
  const int P = mm / 2;
  double* a = smem;                                  // packed m x m upper
  double* rc = a + (m * (m + 1) / 2);                // P cosines
  double* rs = rc + P;                               // P sines
  int* rp = reinterpret_cast<int*>(rs + P);          // P first index
  int* rq = rp + P;                                  // P second index
  __shared__ double red[EIG_NT / 64];
  __shared__ int nrot_s;
  __shared__ double fro_s;
  const int t = threadIdx.x;

  // load and symmetrise (average both triangles so a slightly asymmetric input is still handled)
  for (int b = t; b < m * (m + 1) / 2; b += EIG_NT) {
    int i, j;
    decode_tri(b, i, j);
    a[b] = 0.5 * (H[i + (long long)j * ldh] + H[j + (long long)i * ldh]);
  }
  __syncthreads();
  double part = 0.0;
  for (int b = t; b < m * (m + 1) / 2; b += EIG_NT) {
    int i, j;
    decode_tri(b, i, j);
    double v = a[b];
    part += (i == j ? 1.0 : 2.0) * v * v;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o);
  if ((t & 63) == 0) red[t >> 6] = part;
  __syncthreads();
  if (t == 0) {
    double sacc = 0.0;
    for (int w = 0; w < EIG_NT / 64; ++w) sacc += red[w];
    fro_s = sqrt(sacc);
  }
  __syncthreads();
  // An off-diagonal entry at or below the rounding level of the matrix (2 eps |H|_F) carries no
  // information: rotating it only reshuffles rounding noise, so sweeps would never end on matrices
  // whose spectrum spans many decades.  Above that, the relative criterion resolves small
  // eigenvalues to their own scale.
  const double abs_floor = 4.5e-16 * fro_s + 1e-300;
  const double rel_tol = 1e-15;
  // unordered slot pairs (i <= j) handled by this thread, packed i | j << 16 (-1: none)
  int bij[JB_MAX];
  {
    const int nblk = P * (P + 1) / 2;
#pragma unroll
    for (int u = 0; u < JB_MAX; ++u) {
      const int b = t + u * EIG_NT;
      int i = 0, j = 0;
      if (b < nblk) decode_tri(b, i, j);
      bij[u] = (b < nblk) ? (i | (j << 16)) : -1;
    }
  }

  int round = 0;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    if (t == 0) nrot_s = 0;
    __syncthreads();
    for (int r = 0; r < mm - 1; ++r, ++round) {
      if (t < P) {
        int p, q;
        rr_pair(r, t, mm, p, q);
        double c = 1.0, sn = 0.0;
        if (q < m) {
          double apq = a[pk(p, q)];
          double app = a[pk(p, p)], aqq = a[pk(q, q)];
          if (fabs(apq) > abs_floor && fabs(apq) > rel_tol * sqrt(fabs(app * aqq))) {
            double theta = (aqq - app) / (2.0 * apq);
            double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            sn = tt * c;
            atomicAdd(&nrot_s, 1);
          }
        }
        rc[t] = c; rs[t] = sn; rp[t] = p; rq[t] = q;
        rot[((size_t)round * P + t) * 2 + 0] = c;
        rot[((size_t)round * P + t) * 2 + 1] = sn;
      }
      __syncthreads();
      // 2x2 block updates A' = Ji^T A Jj over unordered pairs of rotation slots (this thread's
      // slot pairs are fixed for the whole run and were decoded once, see bij)
#pragma unroll
      for (int u = 0; u < JB_MAX; ++u) {
        if (bij[u] < 0) continue;
        const int i = bij[u] & 0xffff, j = bij[u] >> 16;
        const double ci = rc[i], si = rs[i], cj = rc[j], sj = rs[j];
        if (si == 0.0 && sj == 0.0) continue;
        const int p = rp[i], q = rq[i];
        if (i == j) {
          if (q >= m) continue;
          double app = a[pk(p, p)], aqq = a[pk(q, q)], apq = a[pk(p, q)];
          double c2 = ci * ci, s2 = si * si, cs = ci * si;
          a[pk(p, p)] = c2 * app - 2.0 * cs * apq + s2 * aqq;
          a[pk(q, q)] = s2 * app + 2.0 * cs * apq + c2 * aqq;
          a[pk(p, q)] = 0.0;
        } else {
          const int r0 = rp[j], s0 = rq[j];
          const bool qv = q < m, sv = s0 < m;
          double xpr = a[pku(p, r0)];
          double xps = sv ? a[pku(p, s0)] : 0.0;
          double xqr = qv ? a[pku(q, r0)] : 0.0;
          double xqs = (qv && sv) ? a[pku(q, s0)] : 0.0;
          double ypr = ci * xpr - si * xqr, yps = ci * xps - si * xqs;
          double yqr = si * xpr + ci * xqr, yqs = si * xps + ci * xqs;
          double zpr = cj * ypr - sj * yps, zps = sj * ypr + cj * yps;
          double zqr = cj * yqr - sj * yqs, zqs = sj * yqr + cj * yqs;
          a[pku(p, r0)] = zpr;
          if (sv) a[pku(p, s0)] = zps;
          if (qv) a[pku(q, r0)] = zqr;
          if (qv && sv) a[pku(q, s0)] = zqs;
        }
      }
      __syncthreads();
    }
    if (nrot_s == 0) { ++sweep; break; }
    __syncthreads();
  }
  for (int i = t; i < m; i += EIG_NT) diag[i] = a[pk(i, i)];
  if (t == 0) {
    *rounds_out = round;
    if (sweeps_out) *sweeps_out = sweep;
  }
}

// Replays the recorded rotations on rows of V (V = product of all rotations,
// starting from the identity).  Each workgroup owns JROWS rows in LDS; rows
// are independent, rotations of one round touch disjoint column pairs.
constexpr int JROWS = 8;
constexpr int JCHUNK = 16;    // rounds of rotation parameters staged in LDS at once
__global__ __launch_bounds__(256) void jacobi_apply_kernel(const double* __restrict__ rot,
                                                           const int* __restrict__ rounds_ptr, int m,
                                                           double* __restrict__ Vrows) {
  extern __shared__ double smem[];
  const int mm = m + (m & 1);
  const int P = mm / 2;
  double* rows = smem;                        // JROWS x m
  double* prm = rows + JROWS * m;             // JCHUNK x P x 2
  const int t = threadIdx.x;
  const int row0 = blockIdx.x * JROWS;
  const int nrows = min(JROWS, m - row0);
  for (int e = t; e < JROWS * m; e += blockDim.x) {
    int rr = e / m, cc = e % m;
    rows[e] = (rr < nrows && row0 + rr == cc) ? 1.0 : 0.0;
  }
  const int R = *rounds_ptr;
  __syncthreads();
  for (int base = 0; base < R; base += JCHUNK) {
    const int nr = min(JCHUNK, R - base);
    for (int e = t; e < nr * P * 2; e += blockDim.x) prm[e] = rot[(size_t)base * P * 2 + e];
    __syncthreads();
    for (int rr = 0; rr < nr; ++rr) {
      const int r_in = (base + rr) % (mm - 1);
      const double* pr = prm + (size_t)rr * P * 2;
      for (int e = t; e < nrows * P; e += blockDim.x) {
        const int row = e / P, slot = e % P;
        const double sn = pr[2 * slot + 1];
        if (sn == 0.0) continue;
        const double c = pr[2 * slot];
        int p, q;
        rr_pair(r_in, slot, mm, p, q);
        double* R_ = rows + row * m;
        double vp = R_[p], vq = R_[q];
        R_[p] = c * vp - sn * vq;
        R_[q] = sn * vp + c * vq;
      }
      __syncthreads();
    }
  }
  for (int e = t; e < nrows * m; e += blockDim.x) {
    int rr = e / m, cc = e % m;
    Vrows[(row0 + rr) + (size_t)m * cc] = rows[rr * m + cc];   // column-major V
  }
}

// sort by |lambda| descending (stable), sign convention: largest |component| of each vector positive
__global__ __launch_bounds__(EIG_NT) void jacobi_sort_kernel(const double* __restrict__ diag,
                                                             const double* __restrict__ Vu, int m,
                                                             double* __restrict__ evals, double* __restrict__ V) {
  for (int i = threadIdx.x; i < m; i += EIG_NT) {
    double li = diag[i];
    double ai = fabs(li);
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      double aj = fabs(diag[j]);
      if (aj > ai || (aj == ai && j < i)) ++rank;
    }
    evals[rank] = li;
    double best = -1.0, sgn = 1.0;
    for (int row = 0; row < m; ++row) {
      double v = Vu[row + (long long)i * m];
      if (fabs(v) > best) { best = fabs(v); sgn = v < 0 ? -1.0 : 1.0; }
    }
    for (int row = 0; row < m; ++row) V[row + (long long)rank * m] = sgn * Vu[row + (long long)i * m];
  }
}

// Cholesky (with diagonal shift) + triangular inverse of an m x m SPD matrix, m <= 192,
// in one workgroup of 32 x 32 threads.  Thread (ty, tx) holds the 6 x 6 elements
// (ty + 32a, tx + 32b) in registers (cyclic distribution, so every step keeps all
// threads busy); each elimination step broadcasts one column through LDS and costs
// one barrier (double-buffered).  L is then parked in LDS (packed lower) and
// B = L^-1 is formed by right-looking forward substitution with the same register
// layout.  The steps run in six 32-column blocks whose block index is a template
// parameter, so every register access is static and blocks that a step cannot
// touch (finished columns, the upper triangle) are not even emitted.
// Output Rinv = L^-T (upper triangular, column-major, zeros below the diagonal) so
// that X = Y Rinv; piv (nullable) receives diag(L).
constexpr int CH_G = 32;                  // thread grid edge
constexpr int CH_R = JACOBI_MAX / CH_G;   // elements per thread per dimension
static_assert(CH_R * CH_G == JACOBI_MAX, "Cholesky tiling");

struct CholShared {
  double* Lp;   // packed lower L, column-major
  double* cb;   // 2 x JACOBI_MAX broadcast buffers
  int* bad;
};

__device__ __forceinline__ int chol_pl(int m, int i, int j) { return j * m - j * (j - 1) / 2 + (i - j); }

// elimination steps jj = 32 BJ .. min(32 BJ + 31, m - 1)
template <int BJ>
__device__ __forceinline__ void chol_block(double (&A)[CH_R][CH_R], int m, const CholShared& sh, int tx, int ty,
                                           int t, double* __restrict__ piv) {
  const int j0 = BJ * CH_G, j1 = min(m, j0 + CH_G);
  for (int jj = j0; jj < j1; ++jj) {
    double* col = sh.cb + (jj & 1) * JACOBI_MAX;
    const int tj = jj - j0;
    if (tx == tj) {
#pragma unroll
      for (int a = BJ; a < CH_R; ++a) {
        const int i = ty + CH_G * a;
        if (i >= jj && i < m) col[i] = A[a][BJ];
      }
    }
    __syncthreads();
    double d = col[jj];
    if (!(d > 0.0)) {
      if (t == 0 && *sh.bad == 0) *sh.bad = jj + 1;
      d = 1e-300;
    }
    const double sd = sqrt(d);
    const double isd = 1.0 / sd;
    if (t == 0 && piv) piv[jj] = sd;
    double li[CH_R], lj[CH_R];
#pragma unroll
    for (int a = BJ; a < CH_R; ++a) {
      const int i = ty + CH_G * a;
      li[a] = (i > jj && i < m) ? col[i] * isd : 0.0;
      const int j = tx + CH_G * a;
      lj[a] = (j > jj && j < m) ? col[j] * isd : 0.0;
    }
    // trailing lower blocks BJ <= b <= a
#pragma unroll
    for (int a = BJ; a < CH_R; ++a)
#pragma unroll
      for (int b = BJ; b <= a; ++b) A[a][b] -= li[a] * lj[b];
    if (tx == tj) {
#pragma unroll
      for (int a = BJ; a < CH_R; ++a) {
        const int i = ty + CH_G * a;
        if (i >= jj) A[a][BJ] = (i == jj) ? sd : li[a];
      }
    }
  }
}

// forward-substitution steps jj = 32 AJ .. min(32 AJ + 31, m - 1) of B = L^-1:
// row jj = B[jj,:] / L[jj,jj] ; rows i > jj: B[i,:] -= L[i,jj] B[jj,:]
template <int AJ>
__device__ __forceinline__ void inv_block(double (&A)[CH_R][CH_R], int m, const CholShared& sh, int tx, int ty) {
  const int j0 = AJ * CH_G, j1 = min(m, j0 + CH_G);
  for (int jj = j0; jj < j1; ++jj) {
    double* row = sh.cb + (jj & 1) * JACOBI_MAX;
    const double ild = 1.0 / sh.Lp[chol_pl(m, jj, jj)];
    if (ty == jj - j0) {
#pragma unroll
      for (int b = 0; b <= AJ; ++b) {   // row jj of B is zero right of the diagonal
        const int j = tx + CH_G * b;
        const double v = A[AJ][b] * ild;
        A[AJ][b] = v;
        if (j < m) row[j] = v;
      }
    }
    __syncthreads();
    double xr[CH_R];
#pragma unroll
    for (int b = 0; b <= AJ; ++b) {
      const int j = tx + CH_G * b;
      xr[b] = (j <= jj) ? row[j] : 0.0;
    }
#pragma unroll
    for (int a = AJ; a < CH_R; ++a) {
      const int i = ty + CH_G * a;
      const double lij = (i > jj && i < m) ? sh.Lp[chol_pl(m, i, jj)] : 0.0;
#pragma unroll
      for (int b = 0; b <= AJ; ++b) A[a][b] -= lij * xr[b];
    }
  }
}

__global__ __launch_bounds__(CH_G * CH_G) void chol_inv_kernel(const double* __restrict__ G, int m,
                                                               double shift, double* __restrict__ Rinv,
                                                               int* info, double* __restrict__ piv) {
  extern __shared__ double smem[];
  __shared__ int bad;
  CholShared sh{smem, smem + (size_t)m * (m + 1) / 2, &bad};
  const int t = threadIdx.x, tx = t % CH_G, ty = t / CH_G;
  double A[CH_R][CH_R];
#pragma unroll
  for (int a = 0; a < CH_R; ++a) {
#pragma unroll
    for (int b = 0; b < CH_R; ++b) {
      const int i = ty + CH_G * a, j = tx + CH_G * b;
      double v = (i == j) ? 1.0 : 0.0;   // padding beyond m: identity, never coupled to the rest
      if (i < m && j < m) v = G[j + (size_t)m * i] + (i == j ? shift : 0.0);
      A[a][b] = v;
    }
  }
  if (t == 0) bad = 0;
  __syncthreads();
  // ---- right-looking Cholesky, six 32-column blocks ----
  chol_block<0>(A, m, sh, tx, ty, t, piv);
  chol_block<1>(A, m, sh, tx, ty, t, piv);
  chol_block<2>(A, m, sh, tx, ty, t, piv);
  chol_block<3>(A, m, sh, tx, ty, t, piv);
  chol_block<4>(A, m, sh, tx, ty, t, piv);
  chol_block<5>(A, m, sh, tx, ty, t, piv);
  // park L (lower part of A) in LDS
#pragma unroll
  for (int a = 0; a < CH_R; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      const int i = ty + CH_G * a, j = tx + CH_G * b;
      if (i < m && j <= i) sh.Lp[chol_pl(m, i, j)] = A[a][b];
    }
  __syncthreads();
  // ---- B = L^-1 ----
#pragma unroll
  for (int a = 0; a < CH_R; ++a)
#pragma unroll
    for (int b = 0; b < CH_R; ++b) A[a][b] = (ty + CH_G * a == tx + CH_G * b) ? 1.0 : 0.0;
  inv_block<0>(A, m, sh, tx, ty);
  inv_block<1>(A, m, sh, tx, ty);
  inv_block<2>(A, m, sh, tx, ty);
  inv_block<3>(A, m, sh, tx, ty);
  inv_block<4>(A, m, sh, tx, ty);
  inv_block<5>(A, m, sh, tx, ty);
  // Rinv(r, c) = B(c, r)
#pragma unroll
  for (int a = 0; a < CH_R; ++a)
#pragma unroll
    for (int b = 0; b < CH_R; ++b) {
      const int i = ty + CH_G * a, j = tx + CH_G * b;
      if (i < m && j < m) Rinv[j + (size_t)m * i] = (j <= i) ? A[a][b] : 0.0;
    }
  if (t == 0 && info) *info = bad;
}


// Rayleigh-Ritz of a nearly diagonal symmetric H (m <= 192, one workgroup) for its first nh
// ("kept") columns: with d = diag(H) and E the off-diagonal part restricted to pairs that involve
// a kept column, when max |E_ij| / |d_i - d_j| over those pairs is <= thresh the eigenpairs are,
// to second order in E,
//   theta_j = d_j + sum_{i != j} E_ij^2 / (d_j - d_i),   v_j = (e_j + sum_{i != j} e_i E_ij / (d_j - d_i)) / norm,
// written sorted by |theta| descending (stable).  Pairs among the further columns (i, j >= nh) are
// left unrotated: those vectors are not converged and only widen the search space, which any
// mixing among them leaves unchanged -- provided the further block cannot hold a direction that
// belongs among the kept ones, which its Gershgorin bound (below the smallest kept |d_i| by 0.1%)
// rules out.  Otherwise *flag = 1 and nothing else is written (the caller falls back to Jacobi).  The vectors are orthonormal to O((E/gap)^2); the subspace iteration's
// residual test sees any remainder.
__global__ __launch_bounds__(EIG_NT) void ritz_perturb_kernel(const double* __restrict__ H, int ldh, int m,
                                                              int nh, double thresh, double* __restrict__ evals,
                                                              double* __restrict__ V, int* __restrict__ flag) {
  __shared__ double dg[JACOBI_MAX], th[JACOBI_MAX], nr[JACOBI_MAX];
  __shared__ int rk[JACOBI_MAX];
  __shared__ double red[EIG_NT / 64];
  __shared__ int bad;
  const int t = threadIdx.x;
  for (int i = t; i < m; i += EIG_NT) dg[i] = H[i + (long long)ldh * i];
  __syncthreads();
  double worst = 0.0;
  for (int e = t; e < m * m; e += EIG_NT) {
    const int i = e % m, j = e / m;
    if (i >= j || i >= nh) continue;   // i < j: pairs with a kept column have i < nh
    const double eij = 0.5 * (H[i + (long long)ldh * j] + H[j + (long long)ldh * i]);
    const double gap = fabs(dg[i] - dg[j]);
    const double r = (eij == 0.0) ? 0.0 : (gap > 0.0 ? fabs(eij) / gap : HUGE_VAL);
    worst = fmax(worst, r);
  }
  // the unrotated further columns must not hide a direction that belongs among the kept ones:
  // Gershgorin bound of the further block below the smallest kept |diagonal| by a margin
  double gmax = 0.0;
  for (int i = nh + t; i < m; i += EIG_NT) {
    double rs = 0.0;
    for (int j = nh; j < m; ++j) rs += fabs(i == j ? dg[i] : 0.5 * (H[i + (long long)ldh * j] + H[j + (long long)ldh * i]));
    gmax = fmax(gmax, rs);
  }
  double hmin = HUGE_VAL;
  for (int i = t; i < nh; i += EIG_NT) hmin = fmin(hmin, fabs(dg[i]));
  for (int o = 32; o > 0; o >>= 1) {
    worst = fmax(worst, __shfl_down(worst, o));
    gmax = fmax(gmax, __shfl_down(gmax, o));
    hmin = fmin(hmin, __shfl_down(hmin, o));
  }
  __shared__ double red2[EIG_NT / 64], red3[EIG_NT / 64];
  if ((t & 63) == 0) { red[t >> 6] = worst; red2[t >> 6] = gmax; red3[t >> 6] = hmin; }
  __syncthreads();
  if (t == 0) {
    double w = 0.0, g = 0.0, hm = HUGE_VAL;
    for (int q = 0; q < EIG_NT / 64; ++q) { w = fmax(w, red[q]); g = fmax(g, red2[q]); hm = fmin(hm, red3[q]); }
    bad = !(w <= thresh) || !(g < (1.0 - 1e-3) * hm);
    *flag = bad;
  }
  __syncthreads();
  if (bad) return;
  auto E = [&](int i, int j) {
    return (i >= nh && j >= nh) ? 0.0 : 0.5 * (H[i + (long long)ldh * j] + H[j + (long long)ldh * i]);
  };
  for (int j = t; j < m; j += EIG_NT) {
    double acc = 0.0, n2 = 1.0;
    for (int i = 0; i < m; ++i) {
      if (i == j) continue;
      const double eij = E(i, j);
      if (eij == 0.0) continue;
      const double kij = eij / (dg[j] - dg[i]);
      acc += eij * kij;
      n2 += kij * kij;
    }
    th[j] = dg[j] + acc;
    nr[j] = 1.0 / sqrt(n2);
  }
  __syncthreads();
  for (int j = t; j < m; j += EIG_NT) {
    const double aj = fabs(th[j]);
    int r = 0;
    for (int l = 0; l < m; ++l) {
      const double al = fabs(th[l]);
      r += (al > aj) || (al == aj && l < j);
    }
    rk[j] = r;
    evals[r] = th[j];
  }
  __syncthreads();
  for (int e = t; e < m * m; e += EIG_NT) {
    const int i = e % m, j = e / m;
    double v;
    if (i == j) {
      v = 1.0;
    } else {
      const double eij = E(i, j);
      v = (eij == 0.0) ? 0.0 : eij / (dg[j] - dg[i]);
    }
    V[i + (long long)m * rk[j]] = v * nr[j];
  }
}
}  // namespace

static size_t jacobi_lds(int m) {
  int mm = m + (m & 1);
  int P = mm / 2;
  return (size_t)m * (m + 1) / 2 * sizeof(double) + 2 * P * sizeof(double) + 2 * P * sizeof(int);
}

size_t jacobi_work_doubles(int m, int max_sweeps) {
  const int mm = m + (m & 1);
  const size_t P = mm / 2;
  // rotation record + unsorted V + diagonal + round counter
  return (size_t)max_sweeps * (mm - 1) * P * 2 + (size_t)m * m + m + 8;
}

void jacobi_eig_ws(const double* H, int ldh, int m, double* evals, double* V, double* Vwork, int max_sweeps,
                   int* sweeps_out, hipStream_t s) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "jacobi_eig: size out of range");
  const int mm = m + (m & 1);
  const size_t P = mm / 2;
  double* rot = Vwork;
  double* Vu = rot + (size_t)max_sweeps * (mm - 1) * P * 2;
  double* diag = Vu + (size_t)m * m;
  int* rounds = reinterpret_cast<int*>(diag + m);
  size_t lds = jacobi_lds(m);
  TNAD_HIP_TRY(hipFuncSetAttribute((const void*)jacobi_rot_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
  hipLaunchKernelGGL(jacobi_rot_kernel, dim3(1), dim3(EIG_NT), lds, s, H, ldh, m, rot, max_sweeps, diag, rounds,
                     sweeps_out);
  TNAD_HIP_TRY(hipGetLastError());
  const size_t lds2 = ((size_t)JROWS * m + (size_t)JCHUNK * P * 2) * sizeof(double);
  hipLaunchKernelGGL(jacobi_apply_kernel, dim3((m + JROWS - 1) / JROWS), dim3(256), lds2, s, rot, rounds, m, Vu);
  TNAD_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(jacobi_sort_kernel, dim3(1), dim3(EIG_NT), 0, s, diag, Vu, m, evals, V);
  TNAD_HIP_TRY(hipGetLastError());
}

void ritz_perturb(const double* H, int ldh, int m, int nh, double thresh, double* evals, double* V, int* flag,
                  hipStream_t s) {
  if (m < 1 || m > JACOBI_MAX || nh < 1 || nh > m) throw Error(TNAD_ERR_ARG, "ritz_perturb: size out of range");
  hipLaunchKernelGGL(ritz_perturb_kernel, dim3(1), dim3(EIG_NT), 0, s, H, ldh, m, nh, thresh, evals, V, flag);
  TNAD_HIP_TRY(hipGetLastError());
}

void chol_inv(const double* G, int m, double shift, double* Rinv, int* info, double* piv, hipStream_t s) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "chol_inv: size out of range");
  size_t lds = ((size_t)m * (m + 1) / 2 + 2 * JACOBI_MAX) * sizeof(double);
  TNAD_HIP_TRY(hipFuncSetAttribute((const void*)chol_inv_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(CH_G * CH_G), lds, s, G, m, shift, Rinv, info, piv);
  TNAD_HIP_TRY(hipGetLastError());
}

}  // namespace tnad
