// ComplexF64 path of the CTMRG energy (src/variationalipeps.jl:28-56 with src/ctmrg.jl:62-153 and its
// reverse sweep) and of the SVD adjoint (src/trg.jl:55-106) on complex matrices.
//
// Storage: a complex tensor of n entries is one device block of 2n doubles, the real plane followed by
// the imaginary plane (both column-major).  The C ABI takes and returns Julia's interleaved ComplexF64
// arrays; the entry points convert at the boundary.
//
// Contractions: every complex einsum is four real strided MFMA contractions (tc) on the planes, with
// the conjugation of either operand folded into the signs.
//
// Hermitian eigenproblems (n <= 96): the real symmetric embedding R = [[Re H, -Im H], [Im H, Re H]]
// (2n x 2n) has every eigenvalue of H twice, with eigenvectors [Re x; Im x] and [-Im x; Re x] for each
// eigenvector x of H.  The FP64 Jacobi solver decomposes R; complex vectors are read off its
// eigenvectors in order of |lambda| with two passes of classical Gram-Schmidt (the partner of each pair
// is i times a vector already taken and is dropped), then each vector's phase is fixed by the rule
// below.  The reference's SVD of a complex Hermitian matrix (src/ctmrg.jl:136) leaves these phases to
// LAPACK, and its complex energy depends on them (the energy contracts l with l, not with conj(l),
// src/variationalipeps.jl:53-54, so it is invariant under real orthogonal changes of the
// environment's gauge only): the phase rule makes the result a defined function of the input.
//
// Phase rule: column x is multiplied by conj(q) / |q| with q = sum_r w_r x_r, w_r = 1 + frac(0.6180339887 r);
// when |q| <= 1e-6 |x| the first entry of largest modulus is made real positive instead.  Real vectors
// stay real.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "api_internal.h"
#include "ctmrg.h"
#include "zc.h"

namespace tnad {
namespace {

constexpr int NT = 256;
inline unsigned nb(size_t n, int per = NT) {
  size_t b = (n + per - 1) / per;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(b, 65535));
}

__device__ __forceinline__ double zblock_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];   // fixed order: every thread gets the same value
  return s;
}

// interleaved (re, im) pairs <-> planar [re | im]
__global__ void deinterleave_kernel(const double* __restrict__ src, size_t n, double* __restrict__ dst) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    dst[e] = src[2 * e];
    dst[n + e] = src[2 * e + 1];
  }
}
__global__ void interleave_kernel(const double* __restrict__ src, size_t n, double* __restrict__ dst) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    dst[2 * e] = src[e];
    dst[2 * e + 1] = src[n + e];
  }
}

// Y = X + X^H (n x n)
__global__ void zherm_add_kernel(const double* __restrict__ X, int n, double* __restrict__ Y) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    const size_t t = j + (size_t)n * i;
    Y[e] = X[e] + X[t];
    Y[nn + e] = X[nn + e] - X[nn + t];
  }
}

// T2[i,j,k] = T1[i,j,k] + conj(T1[k,j,i])   (src/ctmrg.jl:146)
__global__ void zedge_sym_kernel(const double* __restrict__ T1, int chi, int D, double* __restrict__ T2) {
  const size_t N = (size_t)chi * D * chi;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < N; e += (size_t)gridDim.x * blockDim.x) {
    const int i = e % chi, j = (e / chi) % D, k = e / ((size_t)chi * D);
    const size_t t = k + (size_t)chi * (j + (size_t)D * i);
    T2[e] = T1[e] + T1[t];
    T2[N + e] = T1[N + e] - T1[N + t];
  }
}

// R = [[Re H, -Im H], [Im H, Re H]]  (2n x 2n, column-major)
__global__ void realify_kernel(const double* __restrict__ H, int n, double* __restrict__ R) {
  const size_t nn = (size_t)n * n;
  const int m = 2 * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)m * m;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % m, c = e / m;
    const int i = r % n, j = c % n;
    const bool lo_r = r < n, lo_c = c < n;
    const double re = H[i + (size_t)n * j], im = H[nn + i + (size_t)n * j];
    double v;
    if (lo_r == lo_c) v = re;
    else if (lo_r) v = -im;   // top right
    else v = im;              // bottom left
    R[e] = v;
  }
}

__device__ __forceinline__ double phase_weight(int r) {
  const double t = 0.6180339887 * (double)r;
  return 1.0 + (t - floor(t));
}

// Complex eigenvectors of H from the eigenpairs (ev2, V2) of its real embedding (2n of them, sorted by
// |lambda| descending): candidates in order, each orthogonalised against the vectors taken so far (two
// passes of classical Gram-Schmidt), taken when at least half of its norm remains, until n are taken;
// then the phase rule.  One workgroup.  U: planar n x n; lam: n.
constexpr int ZX_NT = 256;
__global__ __launch_bounds__(ZX_NT) void zeig_extract_kernel(const double* __restrict__ V2,
                                                              const double* __restrict__ ev2, int n,
                                                              double* __restrict__ U, double* __restrict__ lam) {
  __shared__ double xr[96], xi[96], cr[96], ci[96];
  __shared__ double red[8];
  __shared__ int taken_s;
  const int t = threadIdx.x;
  const int m = 2 * n;
  const size_t nn = (size_t)n * n;
  if (t == 0) taken_s = 0;
  __syncthreads();
  for (int cand = 0; cand < m; ++cand) {
    const int taken = taken_s;
    if (taken >= n) break;
    for (int r = t; r < n; r += ZX_NT) {
      xr[r] = V2[r + (size_t)m * cand];
      xi[r] = V2[n + r + (size_t)m * cand];
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      // c_a = u_a^H x
      for (int a = t; a < taken; a += ZX_NT) {
        double sr = 0.0, si = 0.0;
        for (int r = 0; r < n; ++r) {
          const double ur = U[r + (size_t)n * a], ui = U[nn + r + (size_t)n * a];
          sr += ur * xr[r] + ui * xi[r];
          si += ur * xi[r] - ui * xr[r];
        }
        cr[a] = sr;
        ci[a] = si;
      }
      __syncthreads();
      // x -= sum_a c_a u_a
      for (int r = t; r < n; r += ZX_NT) {
        double vr = xr[r], vi = xi[r];
        for (int a = 0; a < taken; ++a) {
          const double ur = U[r + (size_t)n * a], ui = U[nn + r + (size_t)n * a];
          vr -= cr[a] * ur - ci[a] * ui;
          vi -= cr[a] * ui + ci[a] * ur;
        }
        xr[r] = vr;
        xi[r] = vi;
      }
      __syncthreads();
    }
    double acc = 0.0;
    for (int r = t; r < n; r += ZX_NT) acc += xr[r] * xr[r] + xi[r] * xi[r];
    const double nrm2 = zblock_sum(acc, red);
    if (nrm2 > 0.25) {
      // phase rule
      double qr = 0.0, qi = 0.0;
      for (int r = t; r < n; r += ZX_NT) {
        const double w = phase_weight(r);
        qr += w * xr[r];
        qi += w * xi[r];
      }
      qr = zblock_sum(qr, red);
      qi = zblock_sum(qi, red);
      const double nrm = sqrt(nrm2);
      if (sqrt(qr * qr + qi * qi) <= 1e-6 * nrm) {
        if (t == 0) {
          int best = 0;
          double bv = -1.0;
          for (int r = 0; r < n; ++r) {
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
      // x <- x conj(q) / (|q| |x|)
      const double pr = qa > 0.0 ? qr / qa : 1.0, pi = qa > 0.0 ? -qi / qa : 0.0;
      for (int r = t; r < n; r += ZX_NT) {
        const double vr = xr[r] / nrm, vi = xi[r] / nrm;
        U[r + (size_t)n * taken] = vr * pr - vi * pi;
        U[nn + r + (size_t)n * taken] = vr * pi + vi * pr;
      }
      if (t == 0) {
        lam[taken] = ev2[cand];
        taken_s = taken + 1;
      }
    }
    __syncthreads();
  }
  // fewer than n taken only through rounding trouble: leave the rest zero
  const int taken = taken_s;
  for (size_t e = (size_t)taken * n + t; e < nn; e += ZX_NT) {
    U[e] = 0.0;
    U[nn + e] = 0.0;
  }
  for (int a = taken + t; a < n; a += ZX_NT) lam[a] = 0.0;
}

// res (n x n) of svd_back (src/trg.jl:76-88) for a Hermitian corner matrix with the cotangent dU on the
// first chi columns only (dS = dV = nothing): UdU = U^H dU (n x chi), J = F o UdU,
//   res = (J + J^H) S + Diagonal(i Im diag(UdU) Sinv)
// F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eta), formed from the rounded squares (no FP contraction)
__global__ void zsvdback_res_kernel(const double* __restrict__ UdU, const double* __restrict__ sv, int n, int chi,
                                    double eta, double* __restrict__ R) {
#pragma clang fp contract(off)
  const size_t nn = (size_t)n * n, nc = (size_t)n * chi;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    auto F = [&](int a, int b) {
      const double sa2 = sv[a] * sv[a], sb2 = sv[b] * sv[b];
      const double d = sb2 - sa2;
      const double d2 = d * d;
      return d / (d2 + eta);
    };
    double vr = 0.0, vi = 0.0;
    if (j < chi) {   // J_ij s_j
      const double f = F(i, j);
      vr += f * UdU[i + (size_t)n * j] * sv[j];
      vi += f * UdU[nc + i + (size_t)n * j] * sv[j];
    }
    if (i < chi) {   // (J^H)_ij s_j = conj(J_ji) s_j
      const double f = F(j, i);
      vr += f * UdU[j + (size_t)n * i] * sv[j];
      vi -= f * UdU[nc + j + (size_t)n * i] * sv[j];
    }
    if (i == j && j < chi) {
      const double s = sv[j];
      const double sinv = s / (s * s + eta);
      vi += UdU[nc + j + (size_t)n * j] * sinv;
    }
    R[e] = vr;
    R[nn + e] = vi;
  }
}

// X[:, j] *= f[j] (real factors), planar n x k
__global__ void zscale_cols_kernel(double* __restrict__ X, int n, int k, const double* __restrict__ f) {
  const size_t N = (size_t)n * k;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < N; e += (size_t)gridDim.x * blockDim.x) {
    const double v = f[e / n];
    X[e] *= v;
    X[N + e] *= v;
  }
}

// out = sum_i x_i y_i (no conjugation), n complex entries; one block; out[0] = re, out[1] = im
__global__ void zdotu_kernel(const double* __restrict__ x, const double* __restrict__ y, size_t n,
                             double* __restrict__ out) {
  __shared__ double red[16];
  double sr = 0.0, si = 0.0;
  for (size_t e = threadIdx.x; e < n; e += blockDim.x) {
    const double ar = x[e], ai = x[n + e], br = y[e], bi = y[n + e];
    sr += ar * br - ai * bi;
    si += ar * bi + ai * br;
  }
  sr = zblock_sum(sr, red);
  si = zblock_sum(si, red);
  if (threadIdx.x == 0) { out[0] = sr; out[1] = si; }
}

// sum of |x|^2 over 2n doubles, one block
__global__ void sumsq_small_kernel(const double* __restrict__ x, size_t n2, double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (size_t e = threadIdx.x; e < n2; e += blockDim.x) s += x[e] * x[e];
  s = zblock_sum(s, red);
  if (threadIdx.x == 0) out[0] = s;
}
// Re <x, y> over 2n doubles (planar: the real dot of the two planes), one block
__global__ void redot_small_kernel(const double* __restrict__ x, const double* __restrict__ y, size_t n2,
                                   double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (size_t e = threadIdx.x; e < n2; e += blockDim.x) s += x[e] * y[e];
  s = zblock_sum(s, red);
  if (threadIdx.x == 0) out[0] = s;
}

// y = x / |x| (nrm2 = |x|^2 on the device); 2n doubles
__global__ void znormalize_kernel(const double* __restrict__ x, size_t n2, const double* __restrict__ nrm2,
                                  double* __restrict__ y) {
  const double nr = sqrt(nrm2[0]);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x)
    y[e] = x[e] / nr;
}
// adjoint of y = x / |x| with the norm adjoint of src/autodiff.jl:23-29:
// xbar = ybar / n + nbar x / (n + eps(0f0)),  nbar = -Re<x, ybar> / n^2
__global__ void znorm_back_kernel(const double* __restrict__ x, const double* __restrict__ ybar, size_t n2,
                                  const double* __restrict__ nrm2, const double* __restrict__ dot,
                                  double* __restrict__ xbar) {
  const double nr = sqrt(nrm2[0]);
  const double nbar = -dot[0] / (nr * nr);
  const double eps0 = 1.401298464324817e-45;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x)
    xbar[e] = ybar[e] / nr + nbar * x[e] / (nr + eps0);
}

// complex bulk symmetrisation (src/ipeps.jl:32-39) and its adjoint, one block, planar in LDS
__constant__ int zc_perm[4][5] = {{0, 3, 2, 1, 4}, {2, 1, 0, 3, 4}, {1, 0, 3, 2, 4}, {3, 2, 1, 0, 4}};
__device__ __forceinline__ int zperm_index(int e, int d, int ph, int which) {
  int idx[5];
  int r = e;
  idx[0] = r % d; r /= d;
  idx[1] = r % d; r /= d;
  idx[2] = r % d; r /= d;
  idx[3] = r % d; r /= d;
  idx[4] = r;
  int j[5];
  for (int k = 0; k < 5; ++k) j[k] = idx[zc_perm[which][k]];
  return j[0] + d * (j[1] + d * (j[2] + d * (j[3] + d * j[4])));
}
__global__ void zsymmetrize_kernel(const double* __restrict__ in, int d, int ph, double* __restrict__ out,
                                   int backward, const double* __restrict__ fwd_in) {
  extern __shared__ double sm[];
  const int N = d * d * d * d * ph;
  double* x = sm;            // 2N
  double* y = sm + 2 * N;    // 2N
  __shared__ double red[16];
  const int t = threadIdx.x;
  auto sweep_fwd = [&]() {
    for (int w = 0; w < 4; ++w) {
      for (int e = t; e < N; e += blockDim.x) {
        const int p = zperm_index(e, d, ph, w);
        y[e] = x[e] + x[p];
        y[N + e] = x[N + e] + x[N + p];
      }
      __syncthreads();
      for (int e = t; e < 2 * N; e += blockDim.x) x[e] = y[e];
      __syncthreads();
    }
  };
  for (int e = t; e < 2 * N; e += blockDim.x) x[e] = backward ? fwd_in[e] : in[e];
  __syncthreads();
  sweep_fwd();
  double acc = 0.0, accd = 0.0;
  for (int e = t; e < 2 * N; e += blockDim.x) {
    acc += x[e] * x[e];
    if (backward) accd += x[e] * in[e];
  }
  const double nrm = sqrt(zblock_sum(acc, red));
  if (!backward) {
    for (int e = t; e < 2 * N; e += blockDim.x) out[e] = x[e] / nrm;
    return;
  }
  const double dot = zblock_sum(accd, red);
  const double nbar = -dot / (nrm * nrm);
  const double eps0 = 1.401298464324817e-45;
  for (int e = t; e < 2 * N; e += blockDim.x) y[e] = in[e] / nrm + nbar * x[e] / (nrm + eps0);
  __syncthreads();
  for (int w = 3; w >= 0; --w) {
    for (int e = t; e < N; e += blockDim.x) {
      const int p = zperm_index(e, d, ph, w);
      x[e] = y[e] + y[p];
      x[N + e] = y[N + e] + y[N + p];
    }
    __syncthreads();
    for (int e = t; e < 2 * N; e += blockDim.x) y[e] = x[e];
    __syncthreads();
  }
  for (int e = t; e < 2 * N; e += blockDim.x) out[e] = y[e];
}

// double layer (src/variationalipeps.jl:30-34): ap[(a,i),(b,j),(c,k),(d,l),x,y] = A[a,b,c,d,x] conj(A[i,j,k,l,y]),
// a = sum_x ap[..., x, x]; planar
__global__ void zdouble_layer_kernel(const double* __restrict__ A, int d, int ph, double* __restrict__ ap,
                                     double* __restrict__ a) {
  const int D = d * d;
  const int D4 = D * D * D * D;
  const int d4 = d * d * d * d;
  const size_t NA = (size_t)d4 * ph, NAP = (size_t)D4 * ph * ph;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < D4; e += gridDim.x * blockDim.x) {
    int r = e;
    const int f0 = r % D; r /= D;
    const int f1 = r % D; r /= D;
    const int f2 = r % D; r /= D;
    const int f3 = r;
    const int ket = (f0 % d) + d * ((f1 % d) + d * ((f2 % d) + d * (f3 % d)));
    const int bra = (f0 / d) + d * ((f1 / d) + d * ((f2 / d) + d * (f3 / d)));
    double trr = 0.0, tri = 0.0;
    for (int y = 0; y < ph; ++y)
      for (int x = 0; x < ph; ++x) {
        const double ar = A[ket + d4 * x], ai = A[NA + ket + d4 * x];
        const double br = A[bra + d4 * y], bi = -A[NA + bra + d4 * y];
        const double vr = ar * br - ai * bi, vi = ar * bi + ai * br;
        if (ap) {
          ap[e + (size_t)D4 * (x + ph * y)] = vr;
          ap[NAP + e + (size_t)D4 * (x + ph * y)] = vi;
        }
        if (x == y) { trr += vr; tri += vi; }
      }
    if (a) { a[e] = trr; a[D4 + e] = tri; }
  }
}

// adjoint: bbar[q,x] = sum_{p,y} (g(ket=q, bra=p, x, y) + conj(g(ket=p, bra=q, y, x))) A[p,y],
// g = apbar + abar (x) delta
__global__ void zdouble_layer_back_kernel(const double* __restrict__ A, int d, int ph,
                                          const double* __restrict__ apbar, const double* __restrict__ abar,
                                          double* __restrict__ bbar) {
  const int D = d * d;
  const int D4 = D * D * D * D;
  const int d4 = d * d * d * d;
  const int N = d4 * ph;
  const size_t NAP = (size_t)D4 * ph * ph;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N; e += gridDim.x * blockDim.x) {
    const int q = e % d4, x = e / d4;
    const int q0 = q % d, q1 = (q / d) % d, q2 = (q / (d * d)) % d, q3 = q / (d * d * d);
    double accr = 0.0, acci = 0.0;
    for (int p = 0; p < d4; ++p) {
      const int p0 = p % d, p1 = (p / d) % d, p2 = (p / (d * d)) % d, p3 = p / (d * d * d);
      const int fqk = (q0 + d * p0) + D * ((q1 + d * p1) + D * ((q2 + d * p2) + D * (q3 + d * p3)));
      const int fpk = (p0 + d * q0) + D * ((p1 + d * q1) + D * ((p2 + d * q2) + D * (p3 + d * q3)));
      for (int y = 0; y < ph; ++y) {
        const size_t o1 = fqk + (size_t)D4 * (x + ph * y), o2 = fpk + (size_t)D4 * (y + ph * x);
        double g1r = apbar ? apbar[o1] : 0.0, g1i = apbar ? apbar[NAP + o1] : 0.0;
        double g2r = apbar ? apbar[o2] : 0.0, g2i = apbar ? apbar[NAP + o2] : 0.0;
        if (x == y && abar) {
          g1r += abar[fqk]; g1i += abar[D4 + fqk];
          g2r += abar[fpk]; g2i += abar[D4 + fpk];
        }
        const double gr = g1r + g2r, gi = g1i - g2i;
        const double ar = A[p + d4 * y], ai = A[N + p + d4 * y];
        accr += gr * ar - gi * ai;
        acci += gr * ai + gi * ar;
      }
    }
    bbar[e] = accr;
    bbar[N + e] = acci;
  }
}

// tr[i,j,k] = sum_a l[i,j,k,a,a] (l: n3 x s x s planar)
__global__ void ztrace_last2_kernel(const double* __restrict__ l, size_t n3, int s, double* __restrict__ tr) {
  const size_t NL = n3 * s * s;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n3; e += (size_t)gridDim.x * blockDim.x) {
    double r = 0.0, i = 0.0;
    for (int a = 0; a < s; ++a) {
      r += l[e + n3 * (a + (size_t)s * a)];
      i += l[NL + e + n3 * (a + (size_t)s * a)];
    }
    tr[e] = r;
    tr[n3 + e] = i;
  }
}
// lbar = cbar1 * conj(L1 + L2) + cbar2 * conj(tr) (x) delta_ab   (planar; cbar1, cbar2 complex scalars on
// the device: sc[0..1] and sc[2..3])
__global__ void zev_lbar_kernel(const double* __restrict__ L1, const double* __restrict__ L2,
                                const double* __restrict__ tr, size_t n3, int s, const double* __restrict__ sc,
                                double* __restrict__ lbar) {
  const size_t NL = n3 * s * s;
  const double ar = sc[0], ai = sc[1], br = sc[2], bi = sc[3];
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < NL; e += (size_t)gridDim.x * blockDim.x) {
    const double lr = L1[e] + L2[e], li = -(L1[NL + e] + L2[NL + e]);   // conj
    double vr = ar * lr - ai * li, vi = ar * li + ai * lr;
    const size_t i3 = e % n3, ab = e / n3;
    const int a = ab % s, b = ab / s;
    if (a == b) {
      const double tr_ = tr[i3], ti = -tr[n3 + i3];   // conj(tr)
      vr += br * tr_ - bi * ti;
      vi += br * ti + bi * tr_;
    }
    lbar[e] = vr;
    lbar[NL + e] = vi;
  }
}

// the scalars of the expectation value's reverse sweep for the loss Re(e / nn):
// sc = [conj(1/nn), conj(-e/nn^2) * 2]; eo = [e re, e im, nn re, nn im]; also writes e/nn to q
__global__ void zev_scalars_kernel(const double* __restrict__ eo, double* __restrict__ sc, double* __restrict__ q) {
  const double er = eo[0], ei = eo[1], nr = eo[2], ni = eo[3];
  const double n2 = nr * nr + ni * ni;
  // 1/nn = conj(nn)/|nn|^2
  const double inr = nr / n2, ini = -ni / n2;
  sc[0] = inr;
  sc[1] = -ini;
  // -e/nn^2 = -e * (1/nn)^2
  const double i2r = inr * inr - ini * ini, i2i = 2 * inr * ini;
  const double tr_ = -(er * i2r - ei * i2i), ti = -(er * i2i + ei * i2r);
  sc[2] = 2.0 * tr_;
  sc[3] = -2.0 * ti;
  q[0] = er * inr - ei * ini;
  q[1] = er * ini + ei * inr;
}

// B = [[0, A], [A^H, 0]]  (N = m + p, planar)
__global__ void zaugment_kernel(const double* __restrict__ A, int m, int p, double* __restrict__ B) {
  const int N = m + p;
  const size_t NN = (size_t)N * N, nA = (size_t)m * p;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < NN; e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % N, c = e / N;
    double vr = 0.0, vi = 0.0;
    if (r < m && c >= m) {            // A[r, c - m]
      const size_t o = r + (size_t)m * (c - m);
      vr = A[o]; vi = A[nA + o];
    } else if (r >= m && c < m) {     // A^H[r - m, c] = conj(A[c, r - m])
      const size_t o = c + (size_t)m * (r - m);
      vr = A[o]; vi = -A[nA + o];
    }
    B[e] = vr;
    B[NN + e] = vi;
  }
}
// U[:, j] = sqrt(2) X[0:m, col_j], V[:, j] = sqrt(2) X[m:, col_j]  (planar)
__global__ void zsplit_kernel(const double* __restrict__ X, int N, int m, const int* __restrict__ cols, int K,
                              double* __restrict__ U, double* __restrict__ V) {
  const int p = N - m;
  const size_t NN = (size_t)N * N, nU = (size_t)m * K, nV = (size_t)p * K;
  const double r2 = 1.4142135623730951;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)N * K;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = e % N, j = e / N;
    const size_t src = r + (size_t)N * cols[j];
    if (r < m) {
      U[r + (size_t)m * j] = r2 * X[src];
      U[nU + r + (size_t)m * j] = r2 * X[NN + src];
    } else {
      V[(r - m) + (size_t)p * j] = r2 * X[src];
      V[nV + (r - m) + (size_t)p * j] = r2 * X[NN + src];
    }
  }
}
// res (K x K) of svd_back (src/trg.jl:76-96) with all three cotangents optional:
//   (J + J^H) S + Diagonal(i Im diag(UdU) Sinv) + S (Kv + Kv^H) + Diagonal(dS),  J = F o UdU, Kv = F o VdV
__global__ void zsvdback_full_res_kernel(const double* __restrict__ UdU, const double* __restrict__ VdV,
                                         const double* __restrict__ dS, const double* __restrict__ sv, int K,
                                         double eta, double* __restrict__ R) {
#pragma clang fp contract(off)
  const size_t KK = (size_t)K * K;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < KK; e += (size_t)gridDim.x * blockDim.x) {
    const int i = e % K, j = e / K;
    const size_t t = j + (size_t)K * i;
    auto F = [&](int a, int b) {
      const double sa2 = sv[a] * sv[a], sb2 = sv[b] * sv[b];
      const double d = sb2 - sa2;
      const double d2 = d * d;
      return d / (d2 + eta);
    };
    const double fij = F(i, j), fji = F(j, i);
    double vr = 0.0, vi = 0.0;
    if (UdU) {   // (J + J^H)_ij s_j
      vr += (fij * UdU[e] + fji * UdU[t]) * sv[j];
      vi += (fij * UdU[KK + e] - fji * UdU[KK + t]) * sv[j];
      if (i == j) vi += UdU[KK + e] * (sv[j] / (sv[j] * sv[j] + eta));
    }
    if (VdV) {   // s_i (Kv + Kv^H)_ij
      vr += sv[i] * (fij * VdV[e] + fji * VdV[t]);
      vi += sv[i] * (fij * VdV[KK + e] - fji * VdV[KK + t]);
    }
    if (dS && i == j) vr += dS[j];
    R[e] = vr;
    R[KK + e] = vi;
  }
}
// sinv[j] = s_j / (s_j^2 + eta)
__global__ void sinv_kernel(const double* __restrict__ sv, int K, double eta, double* __restrict__ out) {
  for (int j = threadIdx.x; j < K; j += blockDim.x) out[j] = sv[j] / (sv[j] * sv[j] + eta);
}

}  // namespace

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
void z_from_interleaved(Ctx* c, const double* src_any, size_t n, double* dst_planar) {
  DevBuf tmp(c, 2 * n * 8);
  copy_in(c, tmp.ptr, src_any, 2 * n * 8);
  hipLaunchKernelGGL(deinterleave_kernel, dim3(nb(n)), dim3(NT), 0, c->stream, tmp.as<double>(), n, dst_planar);
  TNAD_HIP_TRY(hipGetLastError());
}
void z_to_interleaved(Ctx* c, const double* src_planar, size_t n, double* dst_any) {
  DevBuf tmp(c, 2 * n * 8);
  hipLaunchKernelGGL(interleave_kernel, dim3(nb(n)), dim3(NT), 0, c->stream, src_planar, n, tmp.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  copy_out(c, dst_any, tmp.ptr, 2 * n * 8);
}

static std::vector<long long> dims_of(const std::string& idx, const ZDims& dm) {
  std::vector<long long> d;
  for (char ch : idx) {
    auto it = dm.find(ch);
    if (it == dm.end()) throw Error(TNAD_ERR_ARG, std::string("zein: no extent for index ") + ch);
    d.push_back(it->second);
  }
  return d;
}
static size_t count_of(const std::vector<long long>& d) {
  size_t n = 1;
  for (long long x : d) n *= (size_t)x;
  return n;
}

void zein_v(Ctx* c, const TView& vA, const double* A, size_t planeA, bool conjA, const TView& vB, const double* B,
            size_t planeB, bool conjB, const TView& vC, double* Cout, size_t planeC, double alpha, double beta) {
  const double sa = conjA ? -1.0 : 1.0, sb = conjB ? -1.0 : 1.0;
  const double *Ar = A, *Ai = A + planeA, *Br = B, *Bi = B + planeB;
  double *Cr = Cout, *Ci = Cout + planeC;
  // (Ar + i sa Ai)(Br + i sb Bi) = (Ar Br - sa sb Ai Bi) + i (sb Ar Bi + sa Ai Br)
  tc(c, Prec::F64, vA, Ar, vB, Br, vC, Cr, alpha, beta);
  tc(c, Prec::F64, vA, Ai, vB, Bi, vC, Cr, -alpha * sa * sb, 1.0);
  tc(c, Prec::F64, vA, Ar, vB, Bi, vC, Ci, alpha * sb, beta);
  tc(c, Prec::F64, vA, Ai, vB, Br, vC, Ci, alpha * sa, 1.0);
}

void zein(Ctx* c, const std::string& spec, const ZDims& dm, const double* A, bool conjA, const double* B, bool conjB,
          double* Cout, double alpha, double beta) {
  const size_t comma = spec.find(','), arrow = spec.find("->");
  if (comma == std::string::npos || arrow == std::string::npos) throw Error(TNAD_ERR_ARG, "zein: bad spec " + spec);
  const std::string ia = spec.substr(0, comma), ib = spec.substr(comma + 1, arrow - comma - 1),
                    ic = spec.substr(arrow + 2);
  const auto da = dims_of(ia, dm), db = dims_of(ib, dm), dc = dims_of(ic, dm);
  zein_v(c, cm(ia, da), A, count_of(da), conjA, cm(ib, db), B, count_of(db), conjB, cm(ic, dc), Cout, count_of(dc),
         alpha, beta);
}

void z_herm_add(Ctx* c, const double* X, int n, double* Y) {
  hipLaunchKernelGGL(zherm_add_kernel, dim3(nb((size_t)n * n)), dim3(NT), 0, c->stream, X, n, Y);
  TNAD_HIP_TRY(hipGetLastError());
}
void z_edge_sym(Ctx* c, const double* T1, int chi, int D, double* T2) {
  hipLaunchKernelGGL(zedge_sym_kernel, dim3(nb((size_t)chi * D * chi)), dim3(NT), 0, c->stream, T1, chi, D, T2);
  TNAD_HIP_TRY(hipGetLastError());
}

void z_heig(Ctx* c, const double* H, int n, double* U, double* lam) {
  if (n < 1 || 2 * n > JACOBI_MAX) throw Error(TNAD_ERR_UNSUPPORTED, "complex eigensolver: n must be 1..96");
  hipStream_t s = c->stream;
  const int m = 2 * n;
  DevBuf R(c, (size_t)m * m * 8), ev2(c, (size_t)m * 8), V2(c, (size_t)m * m * 8), work(c, syev_dc_work_doubles(m) * 8);
  hipLaunchKernelGGL(realify_kernel, dim3(nb((size_t)m * m)), dim3(NT), 0, s, H, n, R.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  syev_dc(R.as<double>(), m, m, ev2.as<double>(), V2.as<double>(), work.as<double>(), s);
  hipLaunchKernelGGL(zeig_extract_kernel, dim3(1), dim3(ZX_NT), 0, s, V2.as<double>(), ev2.as<double>(), n, U, lam);
  TNAD_HIP_TRY(hipGetLastError());
}

void z_svdback_res(Ctx* c, const double* UdU, const double* VdV, const double* dS, const double* sv, int K, double* R) {
  hipLaunchKernelGGL(zsvdback_full_res_kernel, dim3(nb((size_t)K * K)), dim3(NT), 0, c->stream, UdU, VdV, dS, sv, K,
                     1e-40, R);
  TNAD_HIP_TRY(hipGetLastError());
}

void z_scale_cols(Ctx* c, double* X, int n, int k, const double* f) {
  hipLaunchKernelGGL(zscale_cols_kernel, dim3(nb((size_t)n * k)), dim3(NT), 0, c->stream, X, n, k, f);
  TNAD_HIP_TRY(hipGetLastError());
}

// ---------------------------------------------------------------------------
// complex CTMRG step and its reverse (full spectrum, chi * D <= 96), restated from the reference
// einsum by einsum with pairwise contractions; tp is materialised (chi D^3 chi entries).
// ---------------------------------------------------------------------------
struct ZStep {
  DevBuf C, T, cp, tp, U, lam, sv, sgn, C2, T2, nrm;   // planar complex unless noted (lam, sv, sgn, nrm real)
};

namespace {
struct ZCtx {
  Ctx* c;
  int chi, D, n;
  ZDims dm;
};

DevBuf zbuf(Ctx* c, size_t n) { return DevBuf(c, 2 * n * 8); }

// sign and |lambda| per column; s0 = |lambda_0| on the host
__global__ void zsv_kernel(const double* __restrict__ lam, int n, double* __restrict__ sv, double* __restrict__ sgn) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sv[i] = fabs(lam[i]);
    sgn[i] = lam[i] < 0.0 ? -1.0 : 1.0;
  }
}

void z_normalize(Ctx* c, const double* x, size_t n, double* y, double* nrm2_dev) {
  hipLaunchKernelGGL(sumsq_small_kernel, dim3(1), dim3(1024), 0, c->stream, x, 2 * n, nrm2_dev);
  hipLaunchKernelGGL(znormalize_kernel, dim3(nb(2 * n)), dim3(NT), 0, c->stream, x, 2 * n, nrm2_dev, y);
  TNAD_HIP_TRY(hipGetLastError());
}
void z_norm_back(Ctx* c, const double* x, const double* ybar, size_t n, const double* nrm2_dev, double* xbar) {
  DevBuf dot(c, 16);
  hipLaunchKernelGGL(redot_small_kernel, dim3(1), dim3(1024), 0, c->stream, x, ybar, 2 * n, dot.as<double>());
  hipLaunchKernelGGL(znorm_back_kernel, dim3(nb(2 * n)), dim3(NT), 0, c->stream, x, ybar, 2 * n, nrm2_dev,
                     dot.as<double>(), xbar);
  TNAD_HIP_TRY(hipGetLastError());
}

// one step (src/ctmrg.jl:126-153); vals (host, n) = s ./ s[1]
void zstep_fwd(ZCtx& z, const double* a, const double* C, const double* T, double* Cout, double* Tout,
               std::vector<double>& vals, ZStep& st) {
  Ctx* c = z.c;
  hipStream_t s = c->stream;
  const int chi = z.chi, D = z.D, n = z.n;
  const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi, nn = (size_t)n * n;
  const ZDims& dm = z.dm;
  st.C = zbuf(c, nC); st.T = zbuf(c, nT);
  k_copy(C, st.C.as<double>(), 2 * nC, s);
  k_copy(T, st.T.as<double>(), 2 * nT, s);
  // cp = ein"ad,iba,dcl,jkcb -> ijlk"(C, T, T, a)    (src/ctmrg.jl:130)
  DevBuf X1 = zbuf(c, nT), X2 = zbuf(c, nn);
  st.cp = zbuf(c, nn);
  zein(c, "iBa,ad->iBd", dm, T, false, C, false, X1.as<double>());
  zein(c, "iBd,dCl->iBCl", dm, X1.as<double>(), false, T, false, X2.as<double>());
  zein(c, "iBCl,JKCB->iJlK", dm, X2.as<double>(), false, a, false, st.cp.as<double>());
  // tp = ein"iam,jkla -> ijklm"(T, a)                 (:131)
  st.tp = zbuf(c, nT * D * D);
  zein(c, "iEm,JKLE->iJKLm", dm, T, false, a, false, st.tp.as<double>());
  // cpmat + cpmat'; u, s, v = svd(cpmat)               (:134-136)
  DevBuf M = zbuf(c, nn);
  z_herm_add(c, st.cp.as<double>(), n, M.as<double>());
  st.U = zbuf(c, nn);
  st.lam = DevBuf(c, (size_t)n * 8);
  st.sv = DevBuf(c, (size_t)n * 8);
  st.sgn = DevBuf(c, (size_t)n * 8);
  z_heig(c, M.as<double>(), n, st.U.as<double>(), st.lam.as<double>());
  hipLaunchKernelGGL(zsv_kernel, dim3(1), dim3(256), 0, s, st.lam.as<double>(), n, st.sv.as<double>(),
                     st.sgn.as<double>());
  // z = reshape(u[:, 1:chi], chi, D, chi): the first chi columns of U, planar n x chi
  DevBuf zz = zbuf(c, (size_t)n * chi);
  k_copy(st.U.as<double>(), zz.as<double>(), (size_t)n * chi, s);
  k_copy(st.U.as<double>() + nn, zz.as<double>() + (size_t)n * chi, (size_t)n * chi, s);
  // C1 = ein"abcd,abi,cdj -> ij"(cp, conj(z), z)        (:139)
  DevBuf W1 = zbuf(c, (size_t)n * chi), C1 = zbuf(c, nC);
  zein(c, "aBcD,cDj->aBj", dm, st.cp.as<double>(), false, zz.as<double>(), false, W1.as<double>());
  zein(c, "aBi,aBj->ij", dm, zz.as<double>(), true, W1.as<double>(), false, C1.as<double>());
  // T1 = ein"abjcd,abi,dck -> ijk"(tp, conj(z), z)      (:140)
  DevBuf W2 = zbuf(c, nT * D), T1 = zbuf(c, nT);
  zein(c, "aBJCd,dCk->aBJk", dm, st.tp.as<double>(), false, zz.as<double>(), false, W2.as<double>());
  zein(c, "aBi,aBJk->iJk", dm, zz.as<double>(), true, W2.as<double>(), false, T1.as<double>());
  // corner += corner'; edge += ein"ijk -> kji"(conj(edge))   (:145-146)
  st.C2 = zbuf(c, nC);
  st.T2 = zbuf(c, nT);
  z_herm_add(c, C1.as<double>(), chi, st.C2.as<double>());
  z_edge_sym(c, T1.as<double>(), chi, D, st.T2.as<double>());
  // normalise (:149-150)
  st.nrm = DevBuf(c, 4 * 8);
  z_normalize(c, st.C2.as<double>(), nC, Cout, st.nrm.as<double>() + 0);
  z_normalize(c, st.T2.as<double>(), nT, Tout, st.nrm.as<double>() + 1);
  // vals = s ./ s[1]                                    (:142)
  std::vector<double> hl(n);
  c->read_async(hl.data(), st.lam.ptr, (size_t)n * 8);
  c->sync();
  vals.assign(n, 0.0);
  const double s0 = std::fabs(hl[0]);
  for (int i = 0; i < n; ++i) {
    vals[i] = std::fabs(hl[i]) / s0;
    if (!(vals[i] == vals[i])) throw Error(TNAD_ERR_NAN, "complex ctmrgstep: NaN in singular values");
  }
}

// reverse of zstep_fwd: cotangents of (Cout, Tout) -> cotangents of C, T (overwritten) and a (accumulated)
void zstep_back(ZCtx& z, const double* a, const ZStep& st, const double* C3bar, const double* T3bar, double* Cbar,
                double* Tbar, double* abar) {
  Ctx* c = z.c;
  hipStream_t s = c->stream;
  const int chi = z.chi, D = z.D, n = z.n;
  const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi, nn = (size_t)n * n, nz = (size_t)n * chi;
  const ZDims& dm = z.dm;
  const double* cp = st.cp.as<double>();
  const double* tp = st.tp.as<double>();
  const double* C = st.C.as<double>();
  const double* T = st.T.as<double>();
  DevBuf zz = zbuf(c, nz);
  k_copy(st.U.as<double>(), zz.as<double>(), nz, s);
  k_copy(st.U.as<double>() + nn, zz.as<double>() + nz, nz, s);
  const double* zp = zz.as<double>();
  // norms (:149-150)
  DevBuf C2bar = zbuf(c, nC), T2bar = zbuf(c, nT), C1bar = zbuf(c, nC), T1bar = zbuf(c, nT);
  z_norm_back(c, st.C2.as<double>(), C3bar, nC, st.nrm.as<double>() + 0, C2bar.as<double>());
  z_norm_back(c, st.T2.as<double>(), T3bar, nT, st.nrm.as<double>() + 1, T2bar.as<double>());
  // C2 = C1 + C1', T2 = T1 + conj(T1^{kji}): the adjoints have the same form
  z_herm_add(c, C2bar.as<double>(), chi, C1bar.as<double>());
  z_edge_sym(c, T2bar.as<double>(), chi, D, T1bar.as<double>());
  // T1 = ein"abjcd,abi,dck -> ijk"(tp, conj(z), z)  as  W2 = tp z ; T1 = conj(z)^T W2
  DevBuf W2 = zbuf(c, nT * D), W2bar = zbuf(c, nT * D), tpbar = zbuf(c, nT * D * D);
  zein(c, "aBJCd,dCk->aBJk", dm, tp, false, zp, false, W2.as<double>());
  // W2bar = ein(T1bar, conj(conj z)) ; zcbar += ein(conj(W2), T1bar)
  DevBuf zbar = zbuf(c, nz), zcbar = zbuf(c, nz);
  zein(c, "iJk,aBi->aBJk", dm, T1bar.as<double>(), false, zp, false, W2bar.as<double>());
  zein(c, "aBJk,iJk->aBi", dm, W2.as<double>(), true, T1bar.as<double>(), false, zcbar.as<double>());
  // W2 = ein(tp, z): tpbar = ein(W2bar, conj z), zbar = ein(conj tp, W2bar)
  zein(c, "aBJk,dCk->aBJCd", dm, W2bar.as<double>(), false, zp, true, tpbar.as<double>());
  zein(c, "aBJCd,aBJk->dCk", dm, tp, true, W2bar.as<double>(), false, zbar.as<double>());
  // C1 = ein"abcd,abi,cdj -> ij"(cp, conj(z), z)  as  W1 = cp z ; C1 = conj(z)^T W1
  DevBuf W1 = zbuf(c, nz), W1bar = zbuf(c, nz), cpbar = zbuf(c, nn);
  zein(c, "aBcD,cDj->aBj", dm, cp, false, zp, false, W1.as<double>());
  zein(c, "ij,aBi->aBj", dm, C1bar.as<double>(), false, zp, false, W1bar.as<double>());
  zein(c, "aBj,ij->aBi", dm, W1.as<double>(), true, C1bar.as<double>(), false, zcbar.as<double>(), 1.0, 1.0);
  zein(c, "aBj,cDj->aBcD", dm, W1bar.as<double>(), false, zp, true, cpbar.as<double>());
  zein(c, "aBcD,aBj->cDj", dm, cp, true, W1bar.as<double>(), false, zbar.as<double>(), 1.0, 1.0);
  // z enters as conj(z) too: zbar += conj(zcbar)
  k_axpby(1.0, zcbar.as<double>(), 1.0, zbar.as<double>(), nz, s);
  k_axpby(-1.0, zcbar.as<double>() + nz, 1.0, zbar.as<double>() + nz, nz, s);
  // svd_back (src/trg.jl:72-106), dU = [zbar 0], dS = dV = nothing; then cpmat += cpmat'
  DevBuf UdU = zbuf(c, nz), R = zbuf(c, nn), UR = zbuf(c, nn), Mbar = zbuf(c, nn);
  zein(c, "rs,rt->st", dm, st.U.as<double>(), true, zbar.as<double>(), false, UdU.as<double>());
  hipLaunchKernelGGL(zsvdback_res_kernel, dim3(nb(nn)), dim3(NT), 0, s, UdU.as<double>(), st.sv.as<double>(), n, chi,
                     1e-40, R.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  // Mbar = U res V^H with V = U sign(lambda): (U res) diag(sign) U^H
  zein(c, "rs,sq->rq", dm, st.U.as<double>(), false, R.as<double>(), false, UR.as<double>());
  z_scale_cols(c, UR.as<double>(), n, n, st.sgn.as<double>());
  zein(c, "rq,pq->rp", dm, UR.as<double>(), false, st.U.as<double>(), true, Mbar.as<double>());
  DevBuf cpmatbar = zbuf(c, nn);
  z_herm_add(c, Mbar.as<double>(), n, cpmatbar.as<double>());
  k_axpby(1.0, cpmatbar.as<double>(), 1.0, cpbar.as<double>(), 2 * nn, s);
  // tp = ein"iam,jkla -> ijklm"(T, a)
  zein(c, "iJKLm,JKLE->iEm", dm, tpbar.as<double>(), false, a, true, Tbar);
  zein(c, "iJKLm,iEm->JKLE", dm, tpbar.as<double>(), false, T, true, abar, 1.0, 1.0);
  // cp = ein"ad,iba,dcl,jkcb -> ijlk"(C, T, T, a) = ((T C) T) a with X1 = T C, X2 = X1 T
  DevBuf X1 = zbuf(c, nT), X2 = zbuf(c, nn), X2bar = zbuf(c, nn), X1bar = zbuf(c, nT);
  zein(c, "iBa,ad->iBd", dm, T, false, C, false, X1.as<double>());
  zein(c, "iBd,dCl->iBCl", dm, X1.as<double>(), false, T, false, X2.as<double>());
  zein(c, "iJlK,JKCB->iBCl", dm, cpbar.as<double>(), false, a, true, X2bar.as<double>());
  zein(c, "iBCl,iJlK->JKCB", dm, X2.as<double>(), true, cpbar.as<double>(), false, abar, 1.0, 1.0);
  zein(c, "iBCl,dCl->iBd", dm, X2bar.as<double>(), false, T, true, X1bar.as<double>());
  zein(c, "iBd,iBCl->dCl", dm, X1.as<double>(), true, X2bar.as<double>(), false, Tbar, 1.0, 1.0);
  zein(c, "iBd,ad->iBa", dm, X1bar.as<double>(), false, C, true, Tbar, 1.0, 1.0);
  zein(c, "iBa,iBd->ad", dm, T, true, X1bar.as<double>(), false, Cbar);
}

// expectationvalue (src/variationalipeps.jl:49-56): e/n (complex) and, if apbar != null, the cotangents of
// Re(e/n) with respect to ap, C and T.  Pairwise as E1 = T C, E2 = E1 T, E3 = C T, E4 = E2 apn, l = E4 E3.
void zexpect(Ctx* c, int chi, int D, int ph, const double* h, const double* ap, const double* C, const double* T,
             double* q_host, double* apbar, double* Cbar, double* Tbar) {
  hipStream_t s = c->stream;
  const long long lc = chi, lD = D, ls = ph, n = lc * lD;
  ZDims dm{{'a', chi}, {'b', chi}, {'e', chi}, {'g', chi}, {'i', chi}, {'k', chi},
           {'c', D}, {'d', D}, {'f', D}, {'j', D}, {'l', ph}, {'m', ph}};
  const size_t nT = (size_t)chi * D * chi, D4 = (size_t)D * D * D * D, nap = D4 * ph * ph;
  DevBuf nap2(c, 16), apn = zbuf(c, nap);
  z_normalize(c, ap, nap, apn.as<double>(), nap2.as<double>());
  const size_t n2 = (size_t)n * n, n4 = n2 * ph * ph, n3 = (size_t)n * chi, nl = n3 * ph * ph;
  DevBuf E1 = zbuf(c, nT), E2 = zbuf(c, n2), E3 = zbuf(c, nT), E4 = zbuf(c, n4), l = zbuf(c, nl);
  zein(c, "ica,ab->icb", dm, T, false, C, false, E1.as<double>());
  zein(c, "icb,bde->icde", dm, E1.as<double>(), false, T, false, E2.as<double>());
  zein(c, "eg,gfk->efk", dm, C, false, T, false, E3.as<double>());
  // E4[i,j,e,f,l,m] = sum_{c,d} E2[i,c,d,e] apn[c,j,f,d,l,m], one contraction per physical pair (l, m)
  const TView vE2 = cm("icde", {lc, lD, lD, lc}), vAp = cm("cjfd", {lD, lD, lD, lD});
  const TView vE4s = with_strides("ijef", {lc, lD, lc, lD}, {1, lc, n, n * lc});
  const TView vE4 = with_strides("ijeflm", {lc, lD, lc, lD, ls, ls}, {1, lc, n, n * lc, n * n, n * n * ls});
  const TView vE3 = cm("efk", {lc, lD, lc}), vl = cm("ijklm", {lc, lD, lc, ls, ls});
  for (int lm = 0; lm < ph * ph; ++lm)
    zein_v(c, vE2, E2.as<double>(), n2, false, vAp, apn.as<double>() + lm * D4, nap, false, vE4s,
           E4.as<double>() + lm * n2, n4);
  // l[i,j,k,l,m] = sum_{e,f} E4[i,j,e,f,l,m] E3[e,f,k]
  zein_v(c, vE4, E4.as<double>(), n4, false, vE3, E3.as<double>(), nT, false, vl, l.as<double>(), nl);
  // e = ein"abcij,abckl,ijkl ->"(l, l, h) = sum l (l h);  nn = sum_ijk tr^2 with tr = ein"ijkaa -> ijk"(l)
  ZDims dh{{'a', chi}, {'b', D}, {'c', chi}, {'i', ph}, {'j', ph}, {'k', ph}, {'l', ph}};
  DevBuf L1 = zbuf(c, nl), L2 = zbuf(c, nl), tr = zbuf(c, n3), eo(c, 4 * 8), sc(c, 4 * 8), qd(c, 2 * 8);
  zein(c, "abckl,ijkl->abcij", dh, l.as<double>(), false, h, false, L1.as<double>());
  hipLaunchKernelGGL(zdotu_kernel, dim3(1), dim3(1024), 0, s, l.as<double>(), L1.as<double>(), nl, eo.as<double>());
  hipLaunchKernelGGL(ztrace_last2_kernel, dim3(nb(n3)), dim3(NT), 0, s, l.as<double>(), n3, ph, tr.as<double>());
  hipLaunchKernelGGL(zdotu_kernel, dim3(1), dim3(1024), 0, s, tr.as<double>(), tr.as<double>(), n3,
                     eo.as<double>() + 2);
  hipLaunchKernelGGL(zev_scalars_kernel, dim3(1), dim3(1), 0, s, eo.as<double>(), sc.as<double>(), qd.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  c->read_async(q_host, qd.ptr, 2 * 8);
  if (!apbar) { c->sync(); return; }
  // lbar = conj(1/nn) conj(dl e) + 2 conj(-e/nn^2) conj(tr) (x) delta
  zein(c, "abcij,ijkl->abckl", dh, l.as<double>(), false, h, false, L2.as<double>());
  DevBuf lbar = zbuf(c, nl);
  hipLaunchKernelGGL(zev_lbar_kernel, dim3(nb(nl)), dim3(NT), 0, s, L1.as<double>(), L2.as<double>(), tr.as<double>(),
                     n3, ph, sc.as<double>(), lbar.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  DevBuf E4bar = zbuf(c, n4), E3bar = zbuf(c, nT), E2bar = zbuf(c, n2), apnbar = zbuf(c, nap), E1bar = zbuf(c, nT);
  // l = E4 E3
  zein_v(c, vl, lbar.as<double>(), nl, false, vE3, E3.as<double>(), nT, true, vE4, E4bar.as<double>(), n4);
  zein_v(c, vE4, E4.as<double>(), n4, true, vl, lbar.as<double>(), nl, false, vE3, E3bar.as<double>(), nT);
  // E4 = E2 apn, per physical pair
  for (int lm = 0; lm < ph * ph; ++lm) {
    zein_v(c, vE4s, E4bar.as<double>() + lm * n2, n4, false, vAp, apn.as<double>() + lm * D4, nap, true, vE2,
           E2bar.as<double>(), n2, 1.0, lm == 0 ? 0.0 : 1.0);
    zein_v(c, vE2, E2.as<double>(), n2, true, vE4s, E4bar.as<double>() + lm * n2, n4, false, vAp,
           apnbar.as<double>() + lm * D4, nap);
  }
  zein(c, "icde,bde->icb", dm, E2bar.as<double>(), false, T, true, E1bar.as<double>());
  zein(c, "icb,icde->bde", dm, E1.as<double>(), true, E2bar.as<double>(), false, Tbar);
  zein(c, "icb,ab->ica", dm, E1bar.as<double>(), false, C, true, Tbar, 1.0, 1.0);
  zein(c, "ica,icb->ab", dm, T, true, E1bar.as<double>(), false, Cbar);
  zein(c, "efk,gfk->eg", dm, E3bar.as<double>(), false, T, true, Cbar, 1.0, 1.0);
  zein(c, "eg,efk->gfk", dm, C, true, E3bar.as<double>(), false, Tbar, 1.0, 1.0);
  z_norm_back(c, ap, apnbar.as<double>(), nap, nap2.as<double>(), apbar);
  c->sync();
}

}  // namespace

// energy (src/variationalipeps.jl:28-40) of a complex bulk, with the gradient of Re(energy)
// (src/variationalipeps.jl:71) when grad_planar != null.  All pointers planar device buffers.
void zenergy_run(Ctx* c, const double* h, const double* bulk, int d, int ph, int chi, double tol, int maxit,
                 double* e_host, double* grad, int& iters) {
  hipStream_t s = c->stream;
  const int D = d * d, n = chi * D;
  if (n > JACOBI_MAX / 2)
    throw Error(TNAD_ERR_UNSUPPORTED, "complex CTMRG: chi * D must be <= 96 (full spectrum of the complex corner "
                                      "matrix through its 192 x 192 real embedding)");
  const size_t nb5 = (size_t)d * d * d * d * ph, nap = (size_t)D * D * D * D * ph * ph, na = (size_t)D * D * D * D;
  const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
  // indexperm_symmetrize (src/ipeps.jl:32-39)
  DevBuf bs = zbuf(c, nb5);
  {
    const size_t lds = 4 * nb5 * sizeof(double);
    if (lds > 64 * 1024) throw Error(TNAD_ERR_SHAPE, "complex symmetrize: bulk tensor too large");
    hipLaunchKernelGGL(zsymmetrize_kernel, dim3(1), dim3(1024), lds, s, bulk, d, ph, bs.as<double>(), 0,
                       (const double*)nullptr);
    TNAD_HIP_TRY(hipGetLastError());
  }
  // double layer (src/variationalipeps.jl:30-34)
  DevBuf ap = zbuf(c, nap), a = zbuf(c, na);
  hipLaunchKernelGGL(zdouble_layer_kernel, dim3(nb(na)), dim3(NT), 0, s, bs.as<double>(), d, ph, ap.as<double>(),
                     a.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  // :raw init (src/ctmrg.jl:74-86), linear in a: the real kernel on each plane
  DevBuf Ca = zbuf(c, nC), Ta = zbuf(c, nT), Cb = zbuf(c, nC), Tb = zbuf(c, nT);
  k_init_raw(a.as<double>(), D, chi, Ca.as<double>(), Ta.as<double>(), s);
  k_init_raw(a.as<double>() + na, D, chi, Ca.as<double>() + nC, Ta.as<double>() + nT, s);
  // index extents of the step's einsums: lower case (but p, q, r, s) chi, upper case D, p q r s n
  ZCtx z{c, chi, D, n, ZDims{}};
  for (char ch = 'a'; ch <= 'm'; ++ch) z.dm[ch] = chi;
  z.dm['t'] = chi;
  for (char ch = 'B'; ch <= 'L'; ++ch) z.dm[ch] = D;
  for (char ch : std::string("pqrs")) z.dm[ch] = n;
  // ctmrg (src/ctmrg.jl:110-117) with StopFunction (src/fixedpoint.jl:11-41): counter from -1
  std::vector<std::unique_ptr<ZStep>> tape;
  std::vector<double> oldvals(n, std::numeric_limits<double>::infinity()), vals = oldvals;
  int counter = -1;
  iters = 0;
  while (true) {
    counter += 1;
    if (counter > maxit) break;
    double d2 = 0.0;
    for (int i = 0; i < n; ++i) {
      const double dv = vals[i] - oldvals[i];
      d2 += dv * dv;
    }
    if (std::sqrt(d2) <= tol) break;
    oldvals = vals;
    tape.emplace_back(new ZStep());
    zstep_fwd(z, a.as<double>(), Ca.as<double>(), Ta.as<double>(), Cb.as<double>(), Tb.as<double>(), vals,
              *tape.back());
    std::swap(Ca, Cb);
    std::swap(Ta, Tb);
    iters += 1;
  }
  // expectationvalue (src/variationalipeps.jl:49-56)
  DevBuf apbar = zbuf(c, nap), Cbar = zbuf(c, nC), Tbar = zbuf(c, nT);
  zexpect(c, chi, D, ph, h, ap.as<double>(), Ca.as<double>(), Ta.as<double>(), e_host,
          grad ? apbar.as<double>() : nullptr, Cbar.as<double>(), Tbar.as<double>());
  if (!grad) return;
  // reverse sweep over every step (the :raw init is @nograd, src/autodiff.jl:5)
  DevBuf abar = zbuf(c, na), Cb2 = zbuf(c, nC), Tb2 = zbuf(c, nT);
  k_fill(abar.as<double>(), 2 * na, 0.0, s);
  for (int i = (int)tape.size() - 1; i >= 0; --i) {
    zstep_back(z, a.as<double>(), *tape[i], Cbar.as<double>(), Tbar.as<double>(), Cb2.as<double>(), Tb2.as<double>(),
               abar.as<double>());
    std::swap(Cbar, Cb2);
    std::swap(Tbar, Tb2);
  }
  // double layer and symmetrisation adjoints
  DevBuf bbar = zbuf(c, nb5);
  hipLaunchKernelGGL(zdouble_layer_back_kernel, dim3(nb(nb5, 64)), dim3(64), 0, s, bs.as<double>(), d, ph,
                     apbar.as<double>(), abar.as<double>(), bbar.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  const size_t lds = 4 * nb5 * sizeof(double);
  hipLaunchKernelGGL(zsymmetrize_kernel, dim3(1), dim3(1024), lds, s, bbar.as<double>(), d, ph, grad, 1, bulk);
  TNAD_HIP_TRY(hipGetLastError());
  c->sync();
}

// ---------------------------------------------------------------------------
// SVD of a complex m x p matrix (m + p <= 96) through the Hermitian [[0, A], [A^H, 0]]: its eigenpairs are
// (+-s_i, [u_i; +-v_i] / sqrt(2)).  U (m x K), S (K), V (p x K), K = min(m, p): the K largest positive
// eigenvalues, planar.  Phases by the rule of z_heig applied to [u; v] together, so A = U S V^H.
// ---------------------------------------------------------------------------
void zsvd_small(Ctx* c, const double* A, int m, int p, double* U, double* S, double* V) {
  const int N = m + p, K = std::min(m, p);
  if (m < 1 || p < 1 || N > JACOBI_MAX / 2)
    throw Error(TNAD_ERR_UNSUPPORTED, "complex svd: m + p must be <= 96");
  hipStream_t s = c->stream;
  DevBuf B = zbuf(c, (size_t)N * N), X = zbuf(c, (size_t)N * N), lam(c, (size_t)N * 8);
  hipLaunchKernelGGL(zaugment_kernel, dim3(nb((size_t)N * N)), dim3(NT), 0, s, A, m, p, B.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  z_heig(c, B.as<double>(), N, X.as<double>(), lam.as<double>());
  std::vector<double> hl(N);
  c->read_async(hl.data(), lam.ptr, (size_t)N * 8);
  c->sync();
  std::vector<int> cols;
  std::vector<double> sv;
  for (int i = 0; i < N && (int)cols.size() < K; ++i)
    if (hl[i] > 0.0) { cols.push_back(i); sv.push_back(hl[i]); }
  if ((int)cols.size() < K) throw Error(TNAD_ERR_NAN, "complex svd: fewer positive singular values than min(m, p)");
  DevBuf dcols(c, (size_t)K * sizeof(int) + 16);
  c->write_async(dcols.ptr, cols.data(), (size_t)K * sizeof(int));
  c->write_async(S, sv.data(), (size_t)K * 8);
  hipLaunchKernelGGL(zsplit_kernel, dim3(nb((size_t)N * K)), dim3(NT), 0, s, X.as<double>(), N, m, dcols.as<int>(), K,
                     U, V);
  TNAD_HIP_TRY(hipGetLastError());
  c->sync();
}

// svd_back (src/trg.jl:72-106) of the thin complex SVD: cotangents dU (m x K), dS (K, real), dV (p x K), each
// nullable, -> dA (m x p); all planar device buffers except dS (real).
void zsvd_back_small(Ctx* c, const double* A, int m, int p, const double* dU, const double* dS, const double* dV,
                     double* dA) {
  const int K = std::min(m, p);
  hipStream_t s = c->stream;
  DevBuf U = zbuf(c, (size_t)m * K), V = zbuf(c, (size_t)p * K), sv(c, (size_t)K * 8), si(c, (size_t)K * 8);
  zsvd_small(c, A, m, p, U.as<double>(), sv.as<double>(), V.as<double>());
  hipLaunchKernelGGL(sinv_kernel, dim3(1), dim3(256), 0, s, sv.as<double>(), K, 1e-40, si.as<double>());
  ZDims dm{{'r', m}, {'q', p}, {'i', K}, {'j', K}};
  DevBuf UdU, VdV, R = zbuf(c, (size_t)K * K), UR = zbuf(c, (size_t)m * K);
  if (dU) {
    UdU = zbuf(c, (size_t)K * K);
    zein(c, "ri,rj->ij", dm, U.as<double>(), true, dU, false, UdU.as<double>());
  }
  if (dV) {
    VdV = zbuf(c, (size_t)K * K);
    zein(c, "qi,qj->ij", dm, V.as<double>(), true, dV, false, VdV.as<double>());
  }
  hipLaunchKernelGGL(zsvdback_full_res_kernel, dim3(nb((size_t)K * K)), dim3(NT), 0, s,
                     dU ? UdU.as<double>() : nullptr, dV ? VdV.as<double>() : nullptr, dS, sv.as<double>(), K, 1e-40,
                     R.as<double>());
  TNAD_HIP_TRY(hipGetLastError());
  // dA = U res V^H
  zein(c, "ri,ij->rj", dm, U.as<double>(), false, R.as<double>(), false, UR.as<double>());
  zein(c, "rj,qj->rq", dm, UR.as<double>(), false, V.as<double>(), true, dA);
  // (dU - U U^H dU) Sinv V^H  when U is not square (src/trg.jl:98-100)
  if (dU && m != K) {
    DevBuf P = zbuf(c, (size_t)m * K);
    k_copy(dU, P.as<double>(), 2 * (size_t)m * K, s);
    zein(c, "ri,ij->rj", dm, U.as<double>(), false, UdU.as<double>(), false, P.as<double>(), -1.0, 1.0);
    z_scale_cols(c, P.as<double>(), m, K, si.as<double>());
    zein(c, "rj,qj->rq", dm, P.as<double>(), false, V.as<double>(), true, dA, 1.0, 1.0);
  }
  // U Sinv (dV^H - dV^H V V^H) = U Sinv (dV - V V^H dV)^H  when V is not square (src/trg.jl:102-104)
  if (dV && p != K) {
    DevBuf Q = zbuf(c, (size_t)p * K), US = zbuf(c, (size_t)m * K);
    k_copy(dV, Q.as<double>(), 2 * (size_t)p * K, s);
    zein(c, "qi,ij->qj", dm, V.as<double>(), false, VdV.as<double>(), false, Q.as<double>(), -1.0, 1.0);
    k_copy(U.as<double>(), US.as<double>(), 2 * (size_t)m * K, s);
    z_scale_cols(c, US.as<double>(), m, K, si.as<double>());
    zein(c, "rj,qj->rq", dm, US.as<double>(), false, Q.as<double>(), true, dA, 1.0, 1.0);
  }
  c->sync();
}

}  // namespace tnad
