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

__global__ __launch_bounds__(EIG_NT) void jacobi_eig_kernel(const double* __restrict__ H, int ldh,
                                                            int m, double* __restrict__ evals,
                                                            double* __restrict__ V,
                                                            double* __restrict__ Vw, int max_sweeps,
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
  for (int e = t; e < m * m; e += EIG_NT) Vw[e] = ((e % m) == (e / m)) ? 1.0 : 0.0;
  __syncthreads();
  // Frobenius norm for the absolute threshold
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
    double s = 0.0;
    for (int w = 0; w < EIG_NT / 64; ++w) s += red[w];
    fro_s = sqrt(s);
  }
  __syncthreads();
  const double abs_floor = 1e-18 * fro_s + 1e-300;
  const double rel_tol = 1e-15;

  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    if (t == 0) nrot_s = 0;
    __syncthreads();
    for (int r = 0; r < mm - 1; ++r) {
      if (t < P) {
        int p, q;
        if (t == 0) { p = r; q = mm - 1; }
        else { p = (r + t) % (mm - 1); q = (r - t + (mm - 1)) % (mm - 1); }
        if (p > q) { int tmp = p; p = q; q = tmp; }
        double c = 1.0, s = 0.0;
        if (q < m) {
          double apq = a[pk(p, q)];
          double app = a[pk(p, p)], aqq = a[pk(q, q)];
          if (fabs(apq) > abs_floor && fabs(apq) > rel_tol * sqrt(fabs(app * aqq))) {
            double theta = (aqq - app) / (2.0 * apq);
            double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            s = tt * c;
            atomicAdd(&nrot_s, 1);
          }
        }
        rc[t] = c; rs[t] = s; rp[t] = p; rq[t] = q;
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
          // left: rows (p,q) by Ji^T ; right: cols (r,s) by Jj
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
      // eigenvector columns: V' = V J for every active slot
      for (int e = t; e < P * m; e += EIG_NT) {
        int slot = e / m, row = e % m;
        double s = rs[slot];
        if (s == 0.0) continue;
        double c = rc[slot];
        int p = rp[slot], q = rq[slot];
        double vp = Vw[row + (long long)p * m], vq = Vw[row + (long long)q * m];
        Vw[row + (long long)p * m] = c * vp - s * vq;
        Vw[row + (long long)q * m] = s * vp + c * vq;
      }
      __threadfence_block();
      __syncthreads();
    }
    __syncthreads();
    if (nrot_s == 0) break;
    __syncthreads();
  }
  if (t == 0 && sweeps_out) *sweeps_out = sweep;
  __syncthreads();
  __threadfence();
  // sort by |lambda| descending (stable), fix sign: largest |component| positive
  for (int i = t; i < m; i += EIG_NT) {
    double li = a[pk(i, i)];
    double ai = fabs(li);
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      double aj = fabs(a[pk(j, j)]);
      if (aj > ai || (aj == ai && j < i)) ++rank;
    }
    evals[rank] = li;
    double best = -1.0, sgn = 1.0;
    for (int row = 0; row < m; ++row) {
      double v = Vw[row + (long long)i * m];
      if (fabs(v) > best) { best = fabs(v); sgn = v < 0 ? -1.0 : 1.0; }
    }
    for (int row = 0; row < m; ++row) V[row + (long long)rank * m] = sgn * Vw[row + (long long)i * m];
  }
}

// Cholesky (with diagonal shift) + in-place triangular inverse.  Output Rinv = L^{-T}
// (upper triangular, full m x m column-major, zeros below the diagonal).
__global__ __launch_bounds__(EIG_NT) void chol_inv_kernel(const double* __restrict__ G, int m,
                                                          double shift, double* __restrict__ Rinv,
                                                          int* info) {
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

void jacobi_eig_ws(const double* H, int ldh, int m, double* evals, double* V, double* Vwork,
                   int max_sweeps, int* sweeps_out, hipStream_t s) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "jacobi_eig: size out of range");
  size_t lds = jacobi_lds(m);
  TNAD_HIP_TRY(hipFuncSetAttribute((const void*)jacobi_eig_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(jacobi_eig_kernel, dim3(1), dim3(EIG_NT), lds, s, H, ldh, m, evals, V, Vwork,
                     max_sweeps, sweeps_out);
  TNAD_HIP_TRY(hipGetLastError());
}

void chol_inv(const double* G, int m, double shift, double* Rinv, int* info, hipStream_t s) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "chol_inv: size out of range");
  size_t lds = ((size_t)m * (m + 1) / 2 + m) * sizeof(double);
  TNAD_HIP_TRY(hipFuncSetAttribute((const void*)chol_inv_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(EIG_NT), lds, s, G, m, shift, Rinv, info);
  TNAD_HIP_TRY(hipGetLastError());
}

}  // namespace tnad
