// ComplexF64 TRG (src/trg.jl:13-44) and its reverse sweep (src/trg.jl:55-106), on planar complex buffers
// (zc.h).  Same structure as the real TRG (trg.hip): max-normalisation, two permutations, two truncated SVD
// splits and the four-tensor contraction per iteration; the trace and its log at the end.
//
// Split: the singular triplets of the complex M (m x mp) are the positive eigenpairs (s_i, [u_i; v_i]/sqrt(2))
// of the Hermitian [[0, M], [M^H, 0]].  Up to N = m + mp = 96 its full spectrum comes from z_heig; beyond,
// the top eigenpairs of its real embedding (2N x 2N; every eigenvalue four times: +-s, each twice) come from
// the subspace eigensolver (topk_eig) and complex vectors are read off them with Gram-Schmidt as in z_heig.
// The triplets' phases follow the phase rule of zc.hip applied to [u; v]; TRG's result does not depend on
// them (every new bond pairs u with v of the same split).
//
// Reverse: the reference's svd_back among the kept triplets (zsvdback_full_res_kernel: the complex
// formula with the i Im diag(U^H dU) Sinv term), and the rest of the spectrum by the resolvent solves of
// trg.hip on the real embedding: (s_j^2 - M M^H) x = P b is, for x = [Re x; Im x], the real system with
// R(M) R(M)^T, R(M) = [[Re M, -Im M], [Im M, Re M]], projected off both [Re u; Im u] and [-Im u; Re u] of
// every kept u.
#include <cmath>
#include <complex>
#include <memory>
#include <vector>

#include "api_internal.h"
#include "ctmrg.h"
#include "resolvent.h"
#include "zc.h"

namespace tnad {
namespace {

constexpr int NT = 256;
inline unsigned nb(size_t n) {
  size_t b = (n + NT - 1) / NT;
  if (b > 65535ull * 8) b = 65535ull * 8;
  return (unsigned)(b ? b : 1);
}

// out[j0,j1,j2,j3] = in[i0,i1,i2,i3], j_k = i_{perm[k]} (Julia permutedims), both planes; back: out[i] += in[j]
__global__ void zpermute4_kernel(const double* __restrict__ in, int d0, int d1, int d2, int d3, int p0, int p1,
                                 int p2, int p3, double* __restrict__ out, int back) {
  const int din[4] = {d0, d1, d2, d3};
  const int p[4] = {p0, p1, p2, p3};
  const int dout[4] = {din[p0], din[p1], din[p2], din[p3]};
  const size_t n = (size_t)d0 * d1 * d2 * d3;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int j[4];
    size_t r = e;
    for (int k = 0; k < 4; ++k) { j[k] = r % dout[k]; r /= dout[k]; }
    int i[4];
    for (int k = 0; k < 4; ++k) i[p[k]] = j[k];
    const size_t src = i[0] + (size_t)d0 * (i[1] + (size_t)d1 * (i[2] + (size_t)d2 * i[3]));
    if (!back) {
      out[e] = in[src];
      out[n + e] = in[n + src];
    } else {
      out[src] += in[e];
      out[n + src] += in[n + e];
    }
  }
}

// one block: max |x| over n complex entries and its first index
__global__ void zmaxabs_kernel(const double* __restrict__ x, size_t n, double* out_val, int* out_idx) {
  __shared__ double bv[16];
  __shared__ int bi[16];
  double best = -1.0;
  int idx = 0x7fffffff;
  for (size_t e = threadIdx.x; e < n; e += blockDim.x) {
    const double v = x[e] * x[e] + x[n + e] * x[n + e];
    if (v > best) { best = v; idx = (int)e; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_down(best, o);
    const int oi = __shfl_down(idx, o);
    if (ob > best || (ob == best && oi < idx)) { best = ob; idx = oi; }
  }
  if ((threadIdx.x & 63) == 0) { bv[threadIdx.x >> 6] = best; bi[threadIdx.x >> 6] = idx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = bv[0];
    int ii = bi[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      if (bv[w] > b || (bv[w] == b && bi[w] < ii)) { b = bv[w]; ii = bi[w]; }
    out_val[0] = sqrt(b);
    out_val[1] = x[ii];
    out_val[2] = x[n + ii];
    *out_idx = ii;
  }
}

// trace = sum_ij a[i,j,i,j] (complex); out[0..1]
__global__ void ztrace_kernel(const double* __restrict__ a, int d0, int d1, double* out) {
  __shared__ double red[2][16];
  const size_t n = (size_t)d0 * d1 * d0 * d1;
  double sr = 0.0, si = 0.0;
  for (int e = threadIdx.x; e < d0 * d1; e += blockDim.x) {
    const int i = e % d0, j = e / d0;
    const size_t o = i + (size_t)d0 * (j + (size_t)d1 * (i + (size_t)d0 * j));
    sr += a[o];
    si += a[n + o];
  }
  for (int o = 32; o > 0; o >>= 1) { sr += __shfl_down(sr, o); si += __shfl_down(si, o); }
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = sr; red[1][threadIdx.x >> 6] = si; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a0 = 0.0, a1 = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { a0 += red[0][w]; a1 += red[1][w]; }
    out[0] = a0;
    out[1] = a1;
  }
}
__global__ void ztrace_back_kernel(double* __restrict__ abar, int d0, int d1, double vr, double vi) {
  const size_t n = (size_t)d0 * d1 * d0 * d1;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d0 * d1; e += gridDim.x * blockDim.x) {
    const int i = e % d0, j = e / d0;
    const size_t o = i + (size_t)d0 * (j + (size_t)d1 * (i + (size_t)d0 * j));
    abar[o] = vr;
    abar[n + o] = vi;
  }
}

// B = [[0, M], [M^H, 0]]  (N = m + mp, planar)
__global__ void zaug_kernel(const double* __restrict__ M, int m, int mp, double* __restrict__ B) {
  const int N = m + mp;
  const size_t NN = (size_t)N * N, nM = (size_t)m * mp;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < NN; e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % N, c = e / N;
    double vr = 0.0, vi = 0.0;
    if (r < m && c >= m) {
      const size_t o = r + (size_t)m * (c - m);
      vr = M[o]; vi = M[nM + o];
    } else if (r >= m && c < m) {
      const size_t o = c + (size_t)m * (r - m);
      vr = M[o]; vi = -M[nM + o];
    }
    B[e] = vr;
    B[NN + e] = vi;
  }
}

// R = [[Re X, -Im X], [Im X, Re X]] for a planar rows x cols X (R: 2 rows x 2 cols, real)
__global__ void realify_rect_kernel(const double* __restrict__ X, int rows, int cols, double* __restrict__ R) {
  const size_t n = (size_t)rows * cols;
  const int R2 = 2 * rows;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < 4 * n; e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % R2, c = e / R2;
    const int i = r % rows, j = c % cols;
    const bool lo_r = r < rows, lo_c = c < cols;
    const double re = X[i + (size_t)rows * j], im = X[n + i + (size_t)rows * j];
    R[e] = (lo_r == lo_c) ? re : (lo_r ? -im : im);
  }
}

// W[:, c0 + j] = [Re X_j; Im X_j], W[:, c0 + ncol + j] = [-Im X_j; Re X_j]  (X planar rows x ld columns, the
// columns j0 .. j0 + ncol - 1 of it; W: 2 rows x anything, column-major)
__global__ void realify_pairs_kernel(const double* __restrict__ X, int rows, size_t plane, int j0, int ncol,
                                     double* __restrict__ W, int c0) {
  const int R2 = 2 * rows;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)R2 * ncol;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % R2, j = e / R2;
    const int i = r % rows;
    const size_t src = i + (size_t)rows * (j0 + j);
    const double re = X[src], im = X[plane + src];
    W[r + (size_t)R2 * (c0 + j)] = r < rows ? re : im;
    W[r + (size_t)R2 * (c0 + ncol + j)] = r < rows ? -im : re;
  }
}

// X_j = x2[0:rows, j] + i x2[rows:2 rows, j] for j < ncol (planar rows x ncol)
__global__ void unrealify_kernel(const double* __restrict__ x2, int rows, int ncol, double* __restrict__ X) {
  const size_t n = (size_t)rows * ncol;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int i = e % rows, j = e / rows;
    X[e] = x2[i + (size_t)2 * rows * j];
    X[n + e] = x2[rows + i + (size_t)2 * rows * j];
  }
}

__device__ __forceinline__ double zt_phase_weight(int r) {
  const double t = 0.6180339887 * (double)r;
  return 1.0 + (t - floor(t));
}

// complex eigenvectors (N-dimensional) from ncand real eigenvectors of the embedding (2N x ncand, ld 2N,
// |lambda| descending): the zeig_extract rule of zc.hip for any N (x in dynamic LDS), at most max_take taken.
__global__ __launch_bounds__(256) void zeig_extract_any_kernel(const double* __restrict__ V2,
                                                                const double* __restrict__ ev2, int N, int ncand,
                                                                int max_take, double* __restrict__ U,
                                                                double* __restrict__ lam, int* __restrict__ taken_out) {
  extern __shared__ double sm[];
  double* xr = sm;
  double* xi = sm + N;
  __shared__ double cr[192], ci[192];
  __shared__ double red[8];
  __shared__ int taken_s;
  const int t = threadIdx.x;
  const size_t ld = 2 * (size_t)N, plane = (size_t)N * max_take;
  auto bsum = [&](double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;
  };
  if (t == 0) taken_s = 0;
  __syncthreads();
  for (int cand = 0; cand < ncand; ++cand) {
    const int taken = taken_s;
    if (taken >= max_take) break;
    for (int r = t; r < N; r += blockDim.x) {
      xr[r] = V2[r + ld * cand];
      xi[r] = V2[N + r + ld * cand];
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      for (int a = t; a < taken; a += blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int r = 0; r < N; ++r) {
          const double ur = U[r + (size_t)N * a], ui = U[plane + r + (size_t)N * a];
          sr += ur * xr[r] + ui * xi[r];
          si += ur * xi[r] - ui * xr[r];
        }
        cr[a] = sr;
        ci[a] = si;
      }
      __syncthreads();
      for (int r = t; r < N; r += blockDim.x) {
        double vr = xr[r], vi = xi[r];
        for (int a = 0; a < taken; ++a) {
          const double ur = U[r + (size_t)N * a], ui = U[plane + r + (size_t)N * a];
          vr -= cr[a] * ur - ci[a] * ui;
          vi -= cr[a] * ui + ci[a] * ur;
        }
        xr[r] = vr;
        xi[r] = vi;
      }
      __syncthreads();
    }
    double acc = 0.0;
    for (int r = t; r < N; r += blockDim.x) acc += xr[r] * xr[r] + xi[r] * xi[r];
    const double nrm2 = bsum(acc);
    if (nrm2 > 0.25) {
      double qr = 0.0, qi = 0.0;
      for (int r = t; r < N; r += blockDim.x) {
        const double w = zt_phase_weight(r);
        qr += w * xr[r];
        qi += w * xi[r];
      }
      qr = bsum(qr);
      qi = bsum(qi);
      const double nrm = sqrt(nrm2);
      if (sqrt(qr * qr + qi * qi) <= 1e-6 * nrm) {
        if (t == 0) {
          int best = 0;
          double bv = -1.0;
          for (int r = 0; r < N; ++r) {
            const double v = xr[r] * xr[r] + xi[r] * xi[r];
            if (v > bv) { bv = v; best = r; }
          }
          cr[0] = xr[best];
          ci[0] = xi[best];
        }
        __syncthreads();
        qr = cr[0];
        qi = ci[0];
        __syncthreads();
      }
      const double qa = sqrt(qr * qr + qi * qi);
      const double pr = qa > 0.0 ? qr / qa : 1.0, pi = qa > 0.0 ? -qi / qa : 0.0;
      for (int r = t; r < N; r += blockDim.x) {
        const double vr = xr[r] / nrm, vi = xi[r] / nrm;
        U[r + (size_t)N * taken] = vr * pr - vi * pi;
        U[plane + r + (size_t)N * taken] = vr * pi + vi * pr;
      }
      if (t == 0) {
        lam[taken] = ev2[cand];
        taken_s = taken + 1;
      }
    }
    __syncthreads();
  }
  if (t == 0) *taken_out = taken_s;
}

// U[:, j] = sqrt(2) X[0:m, col_j], V[:, j] = sqrt(2) X[m:, col_j]   (X planar N x ldc columns)
__global__ void zsplit_cols_kernel(const double* __restrict__ X, int N, size_t xplane, int m,
                                   const int* __restrict__ cols, int K, double* __restrict__ U,
                                   double* __restrict__ V) {
  const int mp = N - m;
  const size_t nU = (size_t)m * K, nV = (size_t)mp * K;
  const double r2 = 1.4142135623730951;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)N * K;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % N, j = e / N;
    const size_t src = r + (size_t)N * cols[j];
    if (r < m) {
      U[r + (size_t)m * j] = r2 * X[src];
      U[nU + r + (size_t)m * j] = r2 * X[xplane + src];
    } else {
      V[(r - m) + (size_t)mp * j] = r2 * X[src];
      V[nV + (r - m) + (size_t)mp * j] = r2 * X[xplane + src];
    }
  }
}

// u_out[:, j] = U[:, j] sqrt(s_j) (m x k), v_out[j, c] = sqrt(s_j) conj(V[c, j]) (k x mp); j < kv, the rest 0
__global__ void zscale_split_kernel(const double* __restrict__ U, const double* __restrict__ V,
                                    const double* __restrict__ sq, int m, int mp, int K, int kv, int k,
                                    double* __restrict__ u_out, double* __restrict__ v_out) {
  const size_t nU = (size_t)m * K, nV = (size_t)mp * K, nu = (size_t)m * k, nv = (size_t)k * mp;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)(m + mp) * kv;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % (m + mp), j = e / (m + mp);
    if (r < m) {
      u_out[r + (size_t)m * j] = U[r + (size_t)m * j] * sq[j];
      u_out[nu + r + (size_t)m * j] = U[nU + r + (size_t)m * j] * sq[j];
    } else {
      const int c = r - m;
      v_out[j + (size_t)k * c] = V[c + (size_t)mp * j] * sq[j];
      v_out[nv + j + (size_t)k * c] = -V[nV + c + (size_t)mp * j] * sq[j];
    }
  }
}

// cotangents of u = U sqrt(s), v = sqrt(s) V^H (kk columns): dU = ubar sqrt(s), dV[c, j] = conj(vbar[j, c]) sqrt(s),
// dS_j = Re(<U_j, ubar_j> + <conj(V_j), vbar[j, :]>) / (2 sqrt(s_j))   (one block per column j)
__global__ void zsvd_cotangent_kernel(const double* __restrict__ ubar, const double* __restrict__ vbar,
                                      const double* __restrict__ U, const double* __restrict__ V,
                                      const double* __restrict__ sq, int m, int mp, int k, int K, int kk,
                                      double* __restrict__ dU, double* __restrict__ dV, double* __restrict__ dS) {
  const int j = blockIdx.x;
  const size_t nub = (size_t)m * k, nvb = (size_t)k * mp, nU = (size_t)m * K, nV = (size_t)mp * K;
  const size_t ndU = (size_t)m * kk, ndV = (size_t)mp * kk;
  __shared__ double red[4];
  double acc = 0.0;
  for (int r = threadIdx.x; r < m + mp; r += blockDim.x) {
    if (r < m) {
      const double br = ubar[r + (size_t)m * j], bi = ubar[nub + r + (size_t)m * j];
      const double ur = U[r + (size_t)m * j], ui = U[nU + r + (size_t)m * j];
      dU[r + (size_t)m * j] = br * sq[j];
      dU[ndU + r + (size_t)m * j] = bi * sq[j];
      acc += br * ur + bi * ui;                 // Re(conj(ubar) U)
    } else {
      const int c = r - m;
      const double br = vbar[j + (size_t)k * c], bi = vbar[nvb + j + (size_t)k * c];
      const double vr = V[c + (size_t)mp * j], vi = V[nV + c + (size_t)mp * j];
      dV[c + (size_t)mp * j] = br * sq[j];
      dV[ndV + c + (size_t)mp * j] = -bi * sq[j];
      acc += br * vr - bi * vi;                 // Re(conj(vbar) conj(V))
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double tsum = red[0] + red[1] + red[2] + red[3];
    dS[j] = sq[j] > 0.0 ? tsum / (2.0 * sq[j]) : 0.0;
  }
}

__global__ void ssq_kernel(const double* __restrict__ sv, int K, double* __restrict__ s2, double* __restrict__ sq,
                           double* __restrict__ sinv) {
  for (int j = threadIdx.x; j < K; j += blockDim.x) {
    s2[j] = sv[j] * sv[j];
    sq[j] = sqrt(sv[j]);
    sinv[j] = sv[j] / (sv[j] * sv[j] + 1e-40);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// truncated complex SVD split
// ---------------------------------------------------------------------------
struct ZSvdTape {
  int m = 0, mp = 0, k = 0, K = 0;
  DevBuf M;                     // planar m x mp
  DevBuf U, V, sv, sq;          // K computed triplets: U planar m x K, V planar mp x K, s, sqrt(s)
};

static void zsvd_split(Ctx* c, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol, double* u_out,
                       double* v_out, ZSvdTape& tp) {
  hipStream_t s = c->stream;
  const int m = d1 * d2, mp = d3 * d4, N = m + mp, len = std::min(m, mp);
  tp.m = m; tp.mp = mp;
  tp.M = DevBuf(c, 2 * (size_t)m * mp * 8);
  k_copy(t, tp.M.as<double>(), 2 * (size_t)m * mp, s);
  DevBuf B(c, 2 * (size_t)N * N * 8);
  hipLaunchKernelGGL(zaug_kernel, dim3(nb((size_t)N * N)), dim3(NT), 0, s, t, m, mp, B.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  // complex eigenpairs of B by |lambda| descending: Xc (planar N x nc), lam (nc)
  DevBuf Xc, lam;
  int nc = 0;
  if (2 * N <= JACOBI_MAX) {
    nc = N;
    Xc = DevBuf(c, 2 * (size_t)N * N * 8);
    lam = DevBuf(c, (size_t)N * 8);
    z_heig(c, B.as<double>(), N, Xc.as<double>(), lam.as<double>());
  } else {
    const int N2 = 2 * N;
    const int want_c = std::min(N, 2 * std::min(dmax, len) + 2);
    int want = std::min(N2, 2 * want_c);
    int keig = std::min(JACOBI_MAX, std::max(want + 24, 2 * want));
    keig = std::min(keig, N2);
    want = std::min(want, keig);
    DevBuf R(c, (size_t)N2 * N2 * 8), X(c, (size_t)N2 * keig * 8), AX(c, (size_t)N2 * keig * 8),
        th(c, (size_t)keig * 8), ntk(c, 16);
    hipLaunchKernelGGL(realify_rect_kernel, dim3(nb((size_t)N * N * 4)), dim3(NT), 0, s, B.as<double>(), N, N,
                       R.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
    keig = topk_eig(c, R.as<double>(), N2, want, keig, nullptr, 300, 1e-13, X.as<double>(), AX.as<double>(),
                    th.as<double>(), nullptr);
    const int max_take = std::min(keig, 192);
    Xc = DevBuf(c, 2 * (size_t)N * max_take * 8);
    lam = DevBuf(c, (size_t)max_take * 8);
    const size_t lds = 2 * (size_t)N * 8;
    if (lds > 64 * 1024) throw Error(TNAD_ERR_UNSUPPORTED, "complex trg: split above 4096 + 4096");
    hipLaunchKernelGGL(zeig_extract_any_kernel, dim3(1), dim3(256), lds, s, X.as<double>(), th.as<double>(), N, keig,
                       max_take, Xc.as<double>(), lam.as<double>(), ntk.as<int>());
    TNAD_HIP_TRY(hipGetLastError());
    c->read_async(&nc, ntk.ptr, sizeof(int));
    c->sync();
    // Xc's imaginary plane sits after N x max_take entries; compact the count to what was taken below
    if (nc < 1) throw Error(TNAD_ERR_NAN, "complex trg: no eigenvectors");
    // keep the planar layout of N x max_take (the split kernel takes the plane size)
    tp.K = 0;
    std::vector<double> hl(nc);
    c->read_async(hl.data(), lam.ptr, (size_t)nc * 8);
    c->sync();
    const double zero_level = 1e-14 * std::fabs(hl[0]);
    std::vector<int> pos;
    std::vector<double> hs;
    for (int i = 0; i < nc && (int)pos.size() < len; ++i)
      if (hl[i] > zero_level) { pos.push_back(i); hs.push_back(hl[i]); }
    const int K = (int)pos.size();
    int above = 0;
    for (int i = 0; i < K; ++i) above += hs[i] > tol;
    const int k = std::min(std::min(above + 1, dmax), len);
    tp.k = k;
    tp.K = K;
    const int Kc = std::max(K, 1);
    DevBuf cols(c, (size_t)Kc * sizeof(int) + 16);
    tp.sv = DevBuf(c, (size_t)Kc * 8);
    tp.sq = DevBuf(c, (size_t)std::max(k, Kc) * 8);
    tp.U = DevBuf(c, 2 * (size_t)m * Kc * 8);
    tp.V = DevBuf(c, 2 * (size_t)mp * Kc * 8);
    std::vector<double> hsq(Kc, 0.0);
    for (int j = 0; j < K; ++j) hsq[j] = std::sqrt(hs[j]);
    k_fill(tp.sq.as<double>(), (size_t)std::max(k, Kc), 0.0, s);
    if (K > 0) {
      c->write_async(cols.ptr, pos.data(), (size_t)K * sizeof(int));
      c->write_async(tp.sv.ptr, hs.data(), (size_t)K * 8);
      c->write_async(tp.sq.ptr, hsq.data(), (size_t)K * 8);
      hipLaunchKernelGGL(zsplit_cols_kernel, dim3(nb((size_t)N * K)), dim3(NT), 0, s, Xc.as<double>(), N,
                         (size_t)N * max_take, m, cols.as<int>(), K, tp.U.as<double>(), tp.V.as<double>());
      TNAD_HIP_TRY(hipGetLastError());
    }
    k_fill(u_out, 2 * (size_t)m * k, 0.0, s);
    k_fill(v_out, 2 * (size_t)k * mp, 0.0, s);
    const int kv = std::min(k, K);
    if (kv > 0) {
      hipLaunchKernelGGL(zscale_split_kernel, dim3(nb((size_t)(m + mp) * kv)), dim3(NT), 0, s, tp.U.as<double>(),
                         tp.V.as<double>(), tp.sq.as<double>(), m, mp, K, kv, k, u_out, v_out);
      TNAD_HIP_TRY(hipGetLastError());
    }
    c->sync();
    return;
  }
  // full spectrum (N <= 96)
  std::vector<double> hl(nc);
  c->read_async(hl.data(), lam.ptr, (size_t)nc * 8);
  c->sync();
  const double zero_level = 1e-14 * std::fabs(hl[0]);
  std::vector<int> pos;
  std::vector<double> hs;
  for (int i = 0; i < nc && (int)pos.size() < len; ++i)
    if (hl[i] > zero_level) { pos.push_back(i); hs.push_back(hl[i]); }
  const int K = (int)pos.size();
  int above = 0;
  for (int i = 0; i < K; ++i) above += hs[i] > tol;
  const int k = std::min(std::min(above + 1, dmax), len);
  tp.k = k;
  tp.K = K;
  const int Kc = std::max(K, 1);
  DevBuf cols(c, (size_t)Kc * sizeof(int) + 16);
  tp.sv = DevBuf(c, (size_t)Kc * 8);
  tp.sq = DevBuf(c, (size_t)std::max(k, Kc) * 8);
  tp.U = DevBuf(c, 2 * (size_t)m * Kc * 8);
  tp.V = DevBuf(c, 2 * (size_t)mp * Kc * 8);
  std::vector<double> hsq(Kc, 0.0);
  for (int j = 0; j < K; ++j) hsq[j] = std::sqrt(hs[j]);
  k_fill(tp.sq.as<double>(), (size_t)std::max(k, Kc), 0.0, s);
  if (K > 0) {
    c->write_async(cols.ptr, pos.data(), (size_t)K * sizeof(int));
    c->write_async(tp.sv.ptr, hs.data(), (size_t)K * 8);
    c->write_async(tp.sq.ptr, hsq.data(), (size_t)K * 8);
    hipLaunchKernelGGL(zsplit_cols_kernel, dim3(nb((size_t)N * K)), dim3(NT), 0, s, Xc.as<double>(), N,
                       (size_t)N * N, m, cols.as<int>(), K, tp.U.as<double>(), tp.V.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
  }
  k_fill(u_out, 2 * (size_t)m * k, 0.0, s);
  k_fill(v_out, 2 * (size_t)k * mp, 0.0, s);
  const int kv = std::min(k, K);
  if (kv > 0) {
    hipLaunchKernelGGL(zscale_split_kernel, dim3(nb((size_t)(m + mp) * kv)), dim3(NT), 0, s, tp.U.as<double>(),
                       tp.V.as<double>(), tp.sq.as<double>(), m, mp, K, kv, k, u_out, v_out);
    TNAD_HIP_TRY(hipGetLastError());
  }
  c->sync();
}

// svd_back (src/trg.jl:72-106) for u = U sqrt(s), v = sqrt(s) V^H: the kept pairs by the reference's complex
// formula; the other singular pairs and the null spaces by the resolvent solves on the real embedding
//   x_j = P_U (s_j^2 - M M^H)^-1 P_U (s_j dU_j + M dV_j),   Mbar += X V^H
//   y_j = P_V (s_j^2 - M^H M)^-1 P_V (M^H dU_j + s_j dV_j),  Mbar += U Y^H
static void zsvd_split_back(Ctx* c, const ZSvdTape& tp, const double* ubar, const double* vbar, double* tbar) {
  hipStream_t s = c->stream;
  const int m = tp.m, mp = tp.mp, k = tp.k, K = tp.K;
  const int kk = std::min(k, K);
  k_fill(tbar, 2 * (size_t)m * mp, 0.0, s);
  if (kk == 0) return;
  const double* U = tp.U.as<double>();
  const double* V = tp.V.as<double>();
  const double* M = tp.M.as<double>();
  const size_t nU = (size_t)m * K, nV = (size_t)mp * K;
  DevBuf dU(c, 2 * (size_t)m * kk * 8), dV(c, 2 * (size_t)mp * kk * 8), dS(c, (size_t)kk * 8);
  hipLaunchKernelGGL(zsvd_cotangent_kernel, dim3(kk), dim3(NT), 0, s, ubar, vbar, U, V, tp.sq.as<double>(), m, mp, k, K,
                     kk, dU.as<double>(), dV.as<double>(), dS.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  const TView vU = cm("ri", {m, kk}), vV = cm("qi", {mp, kk}), vdU = cm("rj", {m, kk}), vdV = cm("qj", {mp, kk});
  DevBuf gu(c, 2 * (size_t)kk * kk * 8), gv(c, 2 * (size_t)kk * kk * 8), R(c, 2 * (size_t)kk * kk * 8),
      UR(c, 2 * (size_t)m * kk * 8);
  // U and V hold K columns: views of their first kk columns (the planes sit nU / nV apart)
  zein_v(c, vU, U, nU, true, vdU, dU.as<double>(), (size_t)m * kk, false, cm("ij", {kk, kk}), gu.as<double>(),
         (size_t)kk * kk);
  zein_v(c, vV, V, nV, true, vdV, dV.as<double>(), (size_t)mp * kk, false, cm("ij", {kk, kk}), gv.as<double>(),
         (size_t)kk * kk);
  DevBuf s2(c, (size_t)K * 8), sq(c, (size_t)K * 8), sinv(c, (size_t)K * 8);
  hipLaunchKernelGGL(ssq_kernel, dim3(1), dim3(256), 0, s, tp.sv.as<double>(), K, s2.as<double>(), sq.as<double>(),
                     sinv.as<double>());
  z_svdback_res(c, gu.as<double>(), gv.as<double>(), dS.as<double>(), tp.sv.as<double>(), kk, R.as<double>());
  // Mbar = U_kk R V_kk^H
  zein_v(c, vU, U, nU, false, cm("ij", {kk, kk}), R.as<double>(), (size_t)kk * kk, false, vdU, UR.as<double>(),
         (size_t)m * kk);
  zein_v(c, vdU, UR.as<double>(), (size_t)m * kk, false, vdV, V, nV, true, cm("rq", {m, mp}), tbar, (size_t)m * mp,
         1.0, 0.0);
  if (m == kk && mp == kk) return;
  // the real embedding of M for the resolvent operators
  DevBuf RM(c, (size_t)4 * m * mp * 8), T1(c, (size_t)2 * std::max(m, mp) * 2 * kk * 8);
  hipLaunchKernelGGL(realify_rect_kernel, dim3(nb((size_t)4 * m * mp)), dim3(NT), 0, s, M, m, mp, RM.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  const int r2 = 2 * kk, k2 = 2 * K;
  // shifts: the kept s^2 twice; preconditioner estimates: the further s^2 twice
  DevBuf sig2(c, (size_t)std::max(r2, 1) * 8), tz2(c, (size_t)std::max(k2 - r2, 1) * 8);
  k_copy(s2.as<double>(), sig2.as<double>(), kk, s);
  k_copy(s2.as<double>(), sig2.as<double>() + kk, kk, s);
  if (K > kk) {
    k_copy(s2.as<double>() + kk, tz2.as<double>(), K - kk, s);
    k_copy(s2.as<double>() + kk, tz2.as<double>() + (K - kk), K - kk, s);
  }
  auto solve_side = [&](int rows, const double* Wc, size_t wplane, const double* bc, size_t bplane, bool left,
                        double* Xout) {
    const int R2 = 2 * rows;
    DevBuf W2(c, (size_t)R2 * k2 * 8), b2(c, (size_t)R2 * r2 * 8), x2(c, (size_t)R2 * r2 * 8);
    hipLaunchKernelGGL(realify_pairs_kernel, dim3(nb((size_t)R2 * kk)), dim3(NT), 0, s, Wc, rows, wplane, 0, kk,
                       W2.as<double>(), 0);
    if (K > kk)
      hipLaunchKernelGGL(realify_pairs_kernel, dim3(nb((size_t)R2 * (K - kk))), dim3(NT), 0, s, Wc, rows, wplane, kk,
                         K - kk, W2.as<double>(), r2);
    hipLaunchKernelGGL(realify_pairs_kernel, dim3(nb((size_t)R2 * kk)), dim3(NT), 0, s, bc, rows, bplane, 0, kk,
                       b2.as<double>(), 0);
    TNAD_HIP_TRY(hipGetLastError());
    const int m2 = 2 * m, mp2 = 2 * mp;
    auto apply = [&](const double* Vb, double* out) {
      if (left) {   // R R^T
        tc_launch(c, Prec::F64, gemm_desc('T', 'N', mp2, r2, m2, m2, m2, mp2), RM.as<double>(), Vb, T1.as<double>(), s);
        tc_launch(c, Prec::F64, gemm_desc('N', 'N', m2, r2, mp2, m2, mp2, m2), RM.as<double>(), T1.as<double>(), out, s);
      } else {      // R^T R
        tc_launch(c, Prec::F64, gemm_desc('N', 'N', m2, r2, mp2, m2, mp2, m2), RM.as<double>(), Vb, T1.as<double>(), s);
        tc_launch(c, Prec::F64, gemm_desc('T', 'N', mp2, r2, m2, m2, m2, mp2), RM.as<double>(), T1.as<double>(), out, s);
      }
    };
    resolvent_pcg(c, R2, r2, k2, W2.as<double>(), sig2.as<double>(), tz2.as<double>(), apply, b2.as<double>(),
                  x2.as<double>(), 1e-13, 200);
    hipLaunchKernelGGL(unrealify_kernel, dim3(nb((size_t)rows * kk)), dim3(NT), 0, s, x2.as<double>(), rows, kk, Xout);
    TNAD_HIP_TRY(hipGetLastError());
  };
  if (m > kk) {
    // b = s dU + M dV
    DevBuf bl(c, 2 * (size_t)m * kk * 8), X(c, 2 * (size_t)m * kk * 8);
    k_copy(dU.as<double>(), bl.as<double>(), 2 * (size_t)m * kk, s);
    z_scale_cols(c, bl.as<double>(), m, kk, tp.sv.as<double>());
    zein_v(c, cm("rq", {m, mp}), M, (size_t)m * mp, false, vdV, dV.as<double>(), (size_t)mp * kk, false, vdU,
           bl.as<double>(), (size_t)m * kk, 1.0, 1.0);
    solve_side(m, U, nU, bl.as<double>(), (size_t)m * kk, true, X.as<double>());
    zein_v(c, vdU, X.as<double>(), (size_t)m * kk, false, vdV, V, nV, true, cm("rq", {m, mp}), tbar, (size_t)m * mp,
           1.0, 1.0);
  }
  if (mp > kk) {
    // b' = M^H dU + s dV
    DevBuf br(c, 2 * (size_t)mp * kk * 8), Y(c, 2 * (size_t)mp * kk * 8);
    k_copy(dV.as<double>(), br.as<double>(), 2 * (size_t)mp * kk, s);
    z_scale_cols(c, br.as<double>(), mp, kk, tp.sv.as<double>());
    zein_v(c, cm("rq", {m, mp}), M, (size_t)m * mp, true, vdU, dU.as<double>(), (size_t)m * kk, false, vdV,
           br.as<double>(), (size_t)mp * kk, 1.0, 1.0);
    solve_side(mp, V, nV, br.as<double>(), (size_t)mp * kk, false, Y.as<double>());
    zein_v(c, vdU, U, nU, false, vdV, Y.as<double>(), (size_t)mp * kk, true, cm("rq", {m, mp}), tbar, (size_t)m * mp,
           1.0, 1.0);
  }
}

// ---------------------------------------------------------------------------
// trg (src/trg.jl:13-30), complex
// ---------------------------------------------------------------------------
struct ZTrgIter {
  int du = 0, dr = 0;
  DevBuf a_in;
  double maxval = 1.0, ph_re = 1.0, ph_im = 0.0;   // |a[idx]| and a[idx] / |a[idx]|
  int idx = 0;
  ZSvdTape s1, s2;
  DevBuf dr_t, ul_t, ld_t, ru_t;
  int k1 = 0, k2 = 0;
};

void ztrg_run(Ctx* c, const double* a_planar, int q, int chi, int niter, double tol, double* lnz_out,
              double* abar_out) {
  hipStream_t s = c->stream;
  const bool grad = abar_out != nullptr;
  int du = q, dr = q;
  const size_t na0 = (size_t)q * q * q * q;
  DevBuf a(c, 2 * na0 * 8);
  k_copy(a_planar, a.as<double>(), 2 * na0, s);
  std::vector<std::unique_ptr<ZTrgIter>> tape;
  double lnz_re = 0.0, lnz_im = 0.0;
  DevBuf mv(c, 32), mi(c, 16);
  for (int it = 1; it <= niter; ++it) {
    const size_t na = (size_t)du * dr * du * dr;
    std::unique_ptr<ZTrgIter> T(new ZTrgIter());
    T->du = du; T->dr = dr;
    // maxval = maximum(abs.(a)); a /= maxval; lnZ += 2^(1-n) log(maxval)     (src/trg.jl:16-18)
    hipLaunchKernelGGL(zmaxabs_kernel, dim3(1), dim3(1024), 0, s, a.as<double>(), na, mv.as<double>(), mi.as<int>());
    TNAD_HIP_TRY(hipGetLastError());
    double hv[3] = {0.0, 0.0, 0.0};
    int hi = 0;
    c->read_async(hv, mv.ptr, 3 * 8);
    c->read_async(&hi, mi.ptr, sizeof(int));
    c->sync();
    if (!(hv[0] > 0.0) || !std::isfinite(hv[0])) throw Error(TNAD_ERR_NAN, "complex trg: tensor vanished or is not finite");
    if (grad) {
      T->a_in = DevBuf(c, 2 * na * 8);
      k_copy(a.as<double>(), T->a_in.as<double>(), 2 * na, s);
    }
    T->maxval = hv[0];
    T->ph_re = hv[1] / hv[0];
    T->ph_im = hv[2] / hv[0];
    T->idx = hi;
    k_axpby(1.0 / hv[0], a.as<double>(), 0.0, a.as<double>(), 2 * na, s);
    lnz_re += std::pow(2.0, 1 - it) * std::log(hv[0]);
    // dr_ul = ein"urdl -> drul"(a) ; ld_ru = ein"urdl -> ldru"(a)
    DevBuf p1(c, 2 * na * 8), p2(c, 2 * na * 8);
    hipLaunchKernelGGL(zpermute4_kernel, dim3(nb(na)), dim3(NT), 0, s, a.as<double>(), du, dr, du, dr, 2, 1, 0, 3,
                       p1.as<double>(), 0);
    hipLaunchKernelGGL(zpermute4_kernel, dim3(nb(na)), dim3(NT), 0, s, a.as<double>(), du, dr, du, dr, 3, 2, 1, 0,
                       p2.as<double>(), 0);
    TNAD_HIP_TRY(hipGetLastError());
    const int kmax1 = std::min(chi, du * dr), kmax2 = std::min(chi, dr * du);
    T->dr_t = DevBuf(c, 2 * (size_t)du * dr * kmax1 * 8 + 16);
    T->ul_t = DevBuf(c, 2 * (size_t)kmax1 * du * dr * 8 + 16);
    T->ld_t = DevBuf(c, 2 * (size_t)dr * du * kmax2 * 8 + 16);
    T->ru_t = DevBuf(c, 2 * (size_t)kmax2 * dr * du * 8 + 16);
    zsvd_split(c, p1.as<double>(), du, dr, du, dr, chi, tol, T->dr_t.as<double>(), T->ul_t.as<double>(), T->s1);
    zsvd_split(c, p2.as<double>(), dr, du, dr, du, chi, tol, T->ld_t.as<double>(), T->ru_t.as<double>(), T->s2);
    const int k1 = T->s1.k, k2 = T->s2.k;
    T->k1 = k1; T->k2 = k2;
    // a = ein"npu,por,dom,lmn -> urdl"(dr, ld, ul, ru)                   (src/trg.jl:25)
    const ZDims dm{{'n', du}, {'p', dr}, {'o', du}, {'m', dr}, {'u', k1}, {'d', k1}, {'r', k2}, {'l', k2}};
    DevBuf Xb(c, 2 * (size_t)du * k1 * du * k2 * 8), Yb(c, 2 * (size_t)k1 * du * k2 * du * 8),
        an(c, 2 * (size_t)k1 * k2 * k1 * k2 * 8);
    zein(c, "npu,por->nuor", dm, T->dr_t.as<double>(), false, T->ld_t.as<double>(), false, Xb.as<double>());
    zein(c, "dom,lmn->doln", dm, T->ul_t.as<double>(), false, T->ru_t.as<double>(), false, Yb.as<double>());
    zein(c, "nuor,doln->urdl", dm, Xb.as<double>(), false, Yb.as<double>(), false, an.as<double>());
    a = std::move(an);
    du = k1;
    dr = k2;
    if (grad) tape.push_back(std::move(T));
  }
  // trace = ein"ijij -> "(a)[] ; lnZ += log(trace) / 2^niter               (src/trg.jl:27-28)
  DevBuf tr(c, 16);
  hipLaunchKernelGGL(ztrace_kernel, dim3(1), dim3(256), 0, s, a.as<double>(), du, dr, tr.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  double htr[2] = {0.0, 0.0};
  c->read_async(htr, tr.ptr, 16);
  c->sync();
  const std::complex<double> trc(htr[0], htr[1]);
  if (!(std::abs(trc) > 0.0) || !std::isfinite(std::abs(trc))) throw Error(TNAD_ERR_NAN, "complex trg: zero trace");
  const std::complex<double> lt = std::log(trc) / std::pow(2.0, niter);
  lnz_re += lt.real();
  lnz_im += lt.imag();
  if (lnz_out) { lnz_out[0] = lnz_re; lnz_out[1] = lnz_im; }
  if (!grad) return;

  // ---- reverse sweep of Re(lnZ) --------------------------------------------
  DevBuf abar(c, 2 * (size_t)du * dr * du * dr * 8);
  k_fill(abar.as<double>(), 2 * (size_t)du * dr * du * dr, 0.0, s);
  const std::complex<double> seed = std::conj(1.0 / (std::pow(2.0, niter) * trc));
  hipLaunchKernelGGL(ztrace_back_kernel, dim3(nb((size_t)du * dr)), dim3(NT), 0, s, abar.as<double>(), du, dr,
                     seed.real(), seed.imag());
  TNAD_HIP_TRY(hipGetLastError());
  for (int it = niter; it >= 1; --it) {
    ZTrgIter& T = *tape[it - 1];
    const int du0 = T.du, dr0 = T.dr, k1 = T.k1, k2 = T.k2;
    const size_t na = (size_t)du0 * dr0 * du0 * dr0;
    const ZDims dm{{'n', du0}, {'p', dr0}, {'o', du0}, {'m', dr0}, {'u', k1}, {'d', k1}, {'r', k2}, {'l', k2}};
    DevBuf Xb(c, 2 * (size_t)du0 * k1 * du0 * k2 * 8), Yb(c, 2 * (size_t)k1 * du0 * k2 * du0 * 8);
    zein(c, "npu,por->nuor", dm, T.dr_t.as<double>(), false, T.ld_t.as<double>(), false, Xb.as<double>());
    zein(c, "dom,lmn->doln", dm, T.ul_t.as<double>(), false, T.ru_t.as<double>(), false, Yb.as<double>());
    DevBuf Xbar(c, 2 * (size_t)du0 * k1 * du0 * k2 * 8), Ybar(c, 2 * (size_t)k1 * du0 * k2 * du0 * 8);
    zein(c, "urdl,doln->nuor", dm, abar.as<double>(), false, Yb.as<double>(), true, Xbar.as<double>());
    zein(c, "nuor,urdl->doln", dm, Xb.as<double>(), true, abar.as<double>(), false, Ybar.as<double>());
    DevBuf drb(c, 2 * (size_t)du0 * dr0 * k1 * 8 + 16), ldb(c, 2 * (size_t)dr0 * du0 * k2 * 8 + 16),
        ulb(c, 2 * (size_t)k1 * du0 * dr0 * 8 + 16), rub(c, 2 * (size_t)k2 * dr0 * du0 * 8 + 16);
    zein(c, "nuor,por->npu", dm, Xbar.as<double>(), false, T.ld_t.as<double>(), true, drb.as<double>());
    zein(c, "npu,nuor->por", dm, T.dr_t.as<double>(), true, Xbar.as<double>(), false, ldb.as<double>());
    zein(c, "doln,lmn->dom", dm, Ybar.as<double>(), false, T.ru_t.as<double>(), true, ulb.as<double>());
    zein(c, "dom,doln->lmn", dm, T.ul_t.as<double>(), true, Ybar.as<double>(), false, rub.as<double>());
    DevBuf t1b(c, 2 * na * 8), t2b(c, 2 * na * 8);
    zsvd_split_back(c, T.s1, drb.as<double>(), ulb.as<double>(), t1b.as<double>());
    zsvd_split_back(c, T.s2, ldb.as<double>(), rub.as<double>(), t2b.as<double>());
    DevBuf anb(c, 2 * na * 8);
    k_fill(anb.as<double>(), 2 * na, 0.0, s);
    hipLaunchKernelGGL(zpermute4_kernel, dim3(nb(na)), dim3(NT), 0, s, t1b.as<double>(), du0, dr0, du0, dr0, 2, 1, 0,
                       3, anb.as<double>(), 1);
    hipLaunchKernelGGL(zpermute4_kernel, dim3(nb(na)), dim3(NT), 0, s, t2b.as<double>(), du0, dr0, du0, dr0, 3, 2, 1,
                       0, anb.as<double>(), 1);
    TNAD_HIP_TRY(hipGetLastError());
    // a_n = a_in / maxval, maxval = |a_in[idx]| ; lnZ += 2^(1-n) log(maxval)
    DevBuf red(c, 300 * 8);
    k_dot(anb.as<double>(), T.a_in.as<double>(), 2 * na, red.as<double>(), s);   // Re <anbar, a_in>
    double dotv = 0.0;
    c->read_async(&dotv, red.ptr, 8);
    c->sync();
    const double mv_ = T.maxval;
    const double mbar = std::pow(2.0, 1 - it) / mv_ - dotv / (mv_ * mv_);
    DevBuf ain(c, 2 * na * 8);
    k_axpby(1.0 / mv_, anb.as<double>(), 0.0, ain.as<double>(), 2 * na, s);
    // |a_in[idx]| sends mbar * a_in[idx] / |a_in[idx]| to that entry
    double e2[2] = {0.0, 0.0};
    c->read_async(&e2[0], ain.as<double>() + T.idx, 8);
    c->read_async(&e2[1], ain.as<double>() + na + T.idx, 8);
    c->sync();
    e2[0] += mbar * T.ph_re;
    e2[1] += mbar * T.ph_im;
    c->write_async(ain.as<double>() + T.idx, &e2[0], 8);
    c->write_async(ain.as<double>() + na + T.idx, &e2[1], 8);
    abar = std::move(ain);
  }
  k_copy(abar.as<double>(), abar_out, 2 * na0, s);
  c->sync();
}

}  // namespace tnad
