// Small dense FP64 kernels for the eigensolver: one workgroup each, the
// working matrix held as a packed triangle in LDS (192x192 FP64 = 148 KB of
// the 160 KB a CU offers).
//
//  * jacobi_eig_kernel : cyclic two-sided Jacobi with the round-robin
//    (circle) ordering, so each round applies m/2 disjoint rotations in
//    parallel; every 2x2 block of the matrix is updated by one thread.  The
//    eigenvector matrix accumulates in global memory.  Used for the full
//    spectrum when n <= 192 and for the Rayleigh-Ritz problem of the block
//    subspace iteration.
//  * chol_inv_kernel   : Cholesky of a (shifted) Gram matrix followed by an
//    in-place triangular inverse, for CholeskyQR orthonormalisation.
#include "tnad_internal.h"

namespace tnad {
namespace {

constexpr int EIG_NT = 1024;

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
  const double abs_floor = 1e-18 * fro_s + 1e-300;
  const double rel_tol = 1e-15;

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
      // 2x2 block updates A' = Ji^T A Jj over unordered pairs of rotation slots
      const int nblk = P * (P + 1) / 2;
      for (int b = t; b < nblk; b += EIG_NT) {
        int i, j;
        decode_tri(b, i, j);
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

// Cholesky (with diagonal shift) + in-place triangular inverse.  Output Rinv = L^{-T}
// (upper triangular, full m x m column-major, zeros below the diagonal).
__global__ __launch_bounds__(EIG_NT) void chol_inv_kernel(const double* __restrict__ G, int m,
                                                          double shift, double* __restrict__ Rinv,
                                                          int* info, double* __restrict__ piv) {
  extern __shared__ double smem[];
  // packed lower, column-major: element (i,j), i >= j at  j*m - j(j-1)/2 + (i-j)
  double* L = smem;
  double* col = L + m * (m + 1) / 2;
  __shared__ int bad;
  const int t = threadIdx.x;
  auto pl = [m](int i, int j) { return j * m - j * (j - 1) / 2 + (i - j); };
  for (int j = 0; j < m; ++j)
    for (int i = j + t; i < m; i += EIG_NT) L[pl(i, j)] = G[i + (long long)j * m] + (i == j ? shift : 0.0);
  if (t == 0) bad = 0;
  __syncthreads();
  // right-looking Cholesky
  for (int j = 0; j < m; ++j) {
    if (t == 0) {
      double d = L[pl(j, j)];
      if (!(d > 0.0)) {
        if (bad == 0) bad = j + 1;
        d = 1e-300;
      }
      L[pl(j, j)] = sqrt(d);
      if (piv) piv[j] = L[pl(j, j)];
    }
    __syncthreads();
    double djj = L[pl(j, j)];
    for (int i = j + 1 + t; i < m; i += EIG_NT) L[pl(i, j)] /= djj;
    __syncthreads();
    int rem = m - j - 1;
    int tot = rem * (rem + 1) / 2;
    for (int b = t; b < tot; b += EIG_NT) {
      // (i, k) with j < k <= i < m
      int ii, kk;
      decode_tri(b, kk, ii);    // kk <= ii
      int i = j + 1 + ii, k = j + 1 + kk;
      L[pl(i, k)] -= L[pl(i, j)] * L[pl(k, j)];
    }
    __syncthreads();
  }
  // in-place inverse, columns from last to first:
  // X[j,j] = 1/L[j,j];  X[i,j] = -X[j,j] * sum_{k=j+1..i} X[i,k] L[k,j]
  for (int j = m - 1; j >= 0; --j) {
    for (int i = j + t; i < m; i += EIG_NT) col[i] = L[pl(i, j)];
    __syncthreads();
    double xjj = 1.0 / col[j];
    for (int i = j + 1 + t; i < m; i += EIG_NT) {
      double acc = 0.0;
      for (int k = j + 1; k <= i; ++k) acc += L[pl(i, k)] * col[k];
      L[pl(i, j)] = -xjj * acc;
    }
    if (t == 0) L[pl(j, j)] = xjj;
    __syncthreads();
  }
  // Rinv = X^T : Rinv[r, c] = X[c, r] for r <= c
  for (int e = t; e < m * m; e += EIG_NT) {
    int r = e % m, c = e / m;
    Rinv[e] = (r <= c) ? L[pl(c, r)] : 0.0;
  }
  if (t == 0 && info) *info = bad;
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

void chol_inv(const double* G, int m, double shift, double* Rinv, int* info, double* piv, hipStream_t s) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "chol_inv: size out of range");
  size_t lds = ((size_t)m * (m + 1) / 2 + m) * sizeof(double);
  TNAD_HIP_TRY(hipFuncSetAttribute((const void*)chol_inv_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(EIG_NT), lds, s, G, m, shift, Rinv, info, piv);
  TNAD_HIP_TRY(hipGetLastError());
}

}  // namespace tnad
