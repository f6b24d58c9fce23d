// Elementwise, reduction and small-tensor kernels of the CTMRG/energy path.
// All of these are HBM- or latency-bound; reductions are two-pass and
// deterministic (fixed grid, fixed tree), so repeated runs agree bit for bit.
#include "tnad_internal.h"

namespace tnad {
namespace {

constexpr int NT = 256;
constexpr int RED_BLOCKS = 256;   // first-pass blocks of every reduction

inline unsigned nblk(size_t n, int per = NT) {
  size_t b = (n + per - 1) / per;
  if (b > 65535ull * 8) b = 65535ull * 8;
  return (unsigned)(b ? b : 1);
}

// Y = X + X^T for an n x n column-major matrix, 32x32 LDS tiles so both the
// read of X and of X^T are coalesced.
__global__ void sym_add_kernel(const double* __restrict__ X, int n, double* __restrict__ Y) {
  __shared__ double tile[32][33];
  int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: ty in 0..7
  // load the transposed block X[by.., bx..] into tile
  for (int r = ty; r < 32; r += 8) {
    int i = by + tx, j = bx + r;          // element X[i, j] with i in rows of block (by), j cols (bx)
    tile[r][tx] = (i < n && j < n) ? X[i + (long long)j * n] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int i = bx + tx, j = by + r;          // output Y[i, j] = X[i, j] + X[j, i]
    if (i < n && j < n) Y[i + (long long)j * n] = X[i + (long long)j * n] + tile[tx][r];
  }
}

template <typename S, typename D>
__global__ void cast_kernel(const S* __restrict__ s, D* __restrict__ d, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    d[i] = (D)s[i];
}

__global__ void fill_kernel(double* x, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = v;
}

__global__ void axpby_kernel(double a, const double* __restrict__ x, double b, double* __restrict__ y,
                             size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

__global__ void scale_by_dev_kernel(double* x, size_t n, const double* nrm) {
  double inv = 1.0 / *nrm;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] *= inv;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;
}

__global__ void dot_partial_kernel(const double* __restrict__ x, const double* __restrict__ y, size_t n,
                                   double* __restrict__ part) {
  __shared__ double red[16];
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    acc += x[i] * (y ? y[i] : x[i]);
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void final_sum_kernel(const double* __restrict__ part, int np, double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) *out = s;
}

__global__ void edge_sym_kernel(const double* __restrict__ T1, int chi, int D, double* __restrict__ T2) {
  size_t n = (size_t)chi * D * chi;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int i = e % chi;
    int j = (e / chi) % D;
    int k = e / ((size_t)chi * D);
    T2[e] = T1[e] + T1[k + (size_t)chi * j + (size_t)chi * D * i];
  }
}

__global__ void col_norms_kernel(const double* __restrict__ X, int n, double* __restrict__ nrm) {
  __shared__ double red[16];
  const double* col = X + (long long)blockIdx.x * n;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += col[i] * col[i];
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) nrm[blockIdx.x] = sqrt(s);
}

__global__ void scale_cols_kernel(double* X, int n, int k, const double* scale, int invert) {
  size_t tot = (size_t)n * k;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    int j = e / n;
    double s = scale[j];
    double f = invert ? (s > 0.0 ? 1.0 / s : 0.0) : s;
    X[e] *= f;
  }
}

__device__ __forceinline__ unsigned hash32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
__device__ __forceinline__ double hash_uniform(unsigned seed, unsigned long long idx) {
  unsigned h = hash32(seed ^ hash32((unsigned)idx ^ hash32((unsigned)(idx >> 32) + 0x9e3779b9U)));
  unsigned h2 = hash32(h + 0x85ebca6bU);
  unsigned long long bits = ((unsigned long long)h << 21) ^ (unsigned long long)h2;
  return ((bits & ((1ull << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);   // (0,1)
}

// One block per column j, after col_norms_kernel: a column whose norm is at most rel times the largest
// is refilled with uniform random values in (-1, 1) (hash of the seed and the element index n j + i) and its
// norm recomputed in col_norms_kernel's order; then the column is scaled to unit norm as scale_cols_kernel
// (invert) does.  nrm keeps the first-pass norms (every block reads them for the threshold).
__global__ void normalize_fill_cols_kernel(double* X, int n, int k, const double* __restrict__ nrm, double rel,
                                           unsigned seed) {
  __shared__ double red[16];
  __shared__ double bc;
  const int j = blockIdx.x;
  double mx = 0.0;
  for (int q = threadIdx.x; q < k; q += blockDim.x) mx = fmax(mx, nrm[q]);
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    bc = m;
  }
  __syncthreads();
  const double thresh = rel * bc;
  double v = nrm[j];
  double* col = X + (size_t)j * n;
  if (v <= thresh) {   // the same decision in every thread of the block
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double r = 2.0 * hash_uniform(seed, (size_t)j * n + i) - 1.0;
      col[i] = r;
      acc += r * r;
    }
    __syncthreads();   // red is reused
    const double ssum = block_sum(acc, red);
    if (threadIdx.x == 0) bc = sqrt(ssum);
    __syncthreads();
    v = bc;
  }
  const double f = v > 0.0 ? 1.0 / v : 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] *= f;
}

__global__ void random_fill_flagged_kernel(double* X, int n, int k, const int* flag, unsigned seed) {
  size_t tot = (size_t)n * k;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    int j = e / n;
    if (flag[j]) X[e] = 2.0 * hash_uniform(seed, e) - 1.0;
  }
}

__global__ void random_matrix_kernel(double* X, size_t n, unsigned seed) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    X[e] = 2.0 * hash_uniform(seed, e) - 1.0;
}

// :raw environment (src/ctmrg.jl:74-86): C[i,j] = sum_kl a[i,j,k,l], T[i,j,k] = sum_l a[i,j,k,l],
// embedded in zeros, truncated to chi.
__global__ void init_raw_kernel(const double* __restrict__ a, int D, int chi, double* __restrict__ C,
                                double* __restrict__ T) {
  size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nC + nT; e += (size_t)gridDim.x * blockDim.x) {
    if (e < nC) {
      int i = e % chi, j = e / chi;
      double v = 0.0;
      if (i < D && j < D)
        for (int l = 0; l < D; ++l)
          for (int k = 0; k < D; ++k) v += a[i + D * (j + D * (k + (size_t)D * l))];
      C[e] = v;
    } else {
      size_t f = e - nC;
      int i = f % chi, j = (f / chi) % D, k = f / ((size_t)chi * D);
      double v = 0.0;
      if (i < D && k < D)
        for (int l = 0; l < D; ++l) v += a[i + D * (j + D * (k + (size_t)D * l))];
      T[f] = v;
    }
  }
}

// ---- bulk symmetrisation (src/ipeps.jl:32-39) ------------------------------
// The four permutations are involutions, so (P x)[i] = x[i o p].
__constant__ int c_perm[4][5] = {{0, 3, 2, 1, 4}, {2, 1, 0, 3, 4}, {1, 0, 3, 2, 4}, {3, 2, 1, 0, 4}};

__device__ __forceinline__ int permuted_index(int e, int d, int ph, int which) {
  int idx[5];
  int r = e;
  idx[0] = r % d; r /= d;
  idx[1] = r % d; r /= d;
  idx[2] = r % d; r /= d;
  idx[3] = r % d; r /= d;
  idx[4] = r;
  int j[5];
  for (int k = 0; k < 5; ++k) j[k] = idx[c_perm[which][k]];
  return j[0] + d * (j[1] + d * (j[2] + d * (j[3] + d * j[4])));
}

// one block; forward: x <- x + P_w x for w = 0..3, then / norm.  backward: transposed order.
__global__ void symmetrize_bulk_kernel(const double* __restrict__ in, int d, int ph,
                                       double* __restrict__ out, int backward,
                                       const double* __restrict__ fwd_in) {
  extern __shared__ double sm[];
  const int N = d * d * d * d * ph;
  double* x = sm;
  double* y = sm + N;
  __shared__ double red[16];
  __shared__ double s_norm, s_dot;
  const int t = threadIdx.x;
  if (!backward) {
    for (int e = t; e < N; e += blockDim.x) x[e] = in[e];
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
      for (int e = t; e < N; e += blockDim.x) y[e] = x[e] + x[permuted_index(e, d, ph, w)];
      __syncthreads();
      for (int e = t; e < N; e += blockDim.x) x[e] = y[e];
      __syncthreads();
    }
    double acc = 0.0;
    for (int e = t; e < N; e += blockDim.x) acc += x[e] * x[e];
    double s = block_sum(acc, red);
    if (t == 0) s_norm = sqrt(s);
    __syncthreads();
    for (int e = t; e < N; e += blockDim.x) out[e] = x[e] / s_norm;
  } else {
    // recompute the unnormalised forward result into x
    for (int e = t; e < N; e += blockDim.x) x[e] = fwd_in[e];
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
      for (int e = t; e < N; e += blockDim.x) y[e] = x[e] + x[permuted_index(e, d, ph, w)];
      __syncthreads();
      for (int e = t; e < N; e += blockDim.x) x[e] = y[e];
      __syncthreads();
    }
    double acc = 0.0, accd = 0.0;
    for (int e = t; e < N; e += blockDim.x) { acc += x[e] * x[e]; accd += x[e] * in[e]; }
    double s = block_sum(acc, red);
    __syncthreads();
    double sd = block_sum(accd, red);
    if (t == 0) { s_norm = sqrt(s); s_dot = sd; }
    __syncthreads();
    const double nrm = s_norm;
    const double nbar = -s_dot / (nrm * nrm);
    const double eps0 = 1.401298464324817e-45;   // eps(0f0)
    for (int e = t; e < N; e += blockDim.x) y[e] = in[e] / nrm + nbar * x[e] / (nrm + eps0);
    __syncthreads();
    for (int w = 3; w >= 0; --w) {
      for (int e = t; e < N; e += blockDim.x) x[e] = y[e] + y[permuted_index(e, d, ph, w)];
      __syncthreads();
      for (int e = t; e < N; e += blockDim.x) y[e] = x[e];
      __syncthreads();
    }
    for (int e = t; e < N; e += blockDim.x) out[e] = y[e];
  }
}

// ---- double layer (src/variationalipeps.jl:30-34) --------------------------
// ap[(a,i),(b,j),(c,k),(d,l),x,y] = A[a,b,c,d,x] * A[i,j,k,l,y] (ket index fastest),
// a = sum_x ap[..., x, x].
__global__ void double_layer_kernel(const double* __restrict__ A, int d, int ph, double* __restrict__ ap,
                                    double* __restrict__ a) {
  const int D = d * d;
  const int D4 = D * D * D * D;
  const int d4 = d * d * d * d;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < D4; e += gridDim.x * blockDim.x) {
    int r = e;
    int f0 = r % D; r /= D;
    int f1 = r % D; r /= D;
    int f2 = r % D; r /= D;
    int f3 = r;
    int ket = (f0 % d) + d * ((f1 % d) + d * ((f2 % d) + d * (f3 % d)));
    int bra = (f0 / d) + d * ((f1 / d) + d * ((f2 / d) + d * (f3 / d)));
    double tr = 0.0;
    for (int y = 0; y < ph; ++y)
      for (int x = 0; x < ph; ++x) {
        double v = A[ket + d4 * x] * A[bra + d4 * y];
        if (ap) ap[e + (size_t)D4 * (x + ph * y)] = v;
        if (x == y) tr += v;
      }
    if (a) a[e] = tr;
  }
}

// bbar[q,x] = sum_{p,y} g(ket=q,bra=p,x,y) A[p,y] + sum_{p,y} g(ket=p,bra=q,y,x) A[p,y]
// with g = apbar + abar (x) delta.
__global__ void double_layer_back_kernel(const double* __restrict__ A, int d, int ph,
                                         const double* __restrict__ apbar, const double* __restrict__ abar,
                                         double* __restrict__ bbar) {
  const int D = d * d;
  const int D4 = D * D * D * D;
  const int d4 = d * d * d * d;
  const int N = d4 * ph;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N; e += gridDim.x * blockDim.x) {
    int q = e % d4, x = e / d4;
    int q0 = q % d, q1 = (q / d) % d, q2 = (q / (d * d)) % d, q3 = q / (d * d * d);
    double acc = 0.0;
    for (int p = 0; p < d4; ++p) {
      int p0 = p % d, p1 = (p / d) % d, p2 = (p / (d * d)) % d, p3 = p / (d * d * d);
      // fused index with q as ket, p as bra
      int fqk = (q0 + d * p0) + D * ((q1 + d * p1) + D * ((q2 + d * p2) + D * (q3 + d * p3)));
      // fused index with p as ket, q as bra
      int fpk = (p0 + d * q0) + D * ((p1 + d * q1) + D * ((p2 + d * q2) + D * (p3 + d * q3)));
      for (int y = 0; y < ph; ++y) {
        double g1 = (apbar ? apbar[fqk + (size_t)D4 * (x + ph * y)] : 0.0) + (x == y && abar ? abar[fqk] : 0.0);
        double g2 = (apbar ? apbar[fpk + (size_t)D4 * (y + ph * x)] : 0.0) + (x == y && abar ? abar[fpk] : 0.0);
        acc += (g1 + g2) * A[p + d4 * y];
      }
    }
    bbar[e] = acc;
  }
}

}  // namespace

void k_sym_add(const double* X, int n, double* Y, hipStream_t s) {
  dim3 g((n + 31) / 32, (n + 31) / 32);
  hipLaunchKernelGGL(sym_add_kernel, g, dim3(256), 0, s, X, n, Y);
  TNAD_HIP_TRY(hipGetLastError());
}
template <typename T> void k_cast(const T* src, double* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL((cast_kernel<T, double>), dim3(nblk(n)), dim3(NT), 0, s, src, dst, n);
  TNAD_HIP_TRY(hipGetLastError());
}
template <typename T> void k_cast_from(const double* src, T* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL((cast_kernel<double, T>), dim3(nblk(n)), dim3(NT), 0, s, src, dst, n);
  TNAD_HIP_TRY(hipGetLastError());
}
template void k_cast<float>(const float*, double*, size_t, hipStream_t);
template void k_cast_from<float>(const double*, float*, size_t, hipStream_t);

void k_fill(double* x, size_t n, double v, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3(nblk(n)), dim3(NT), 0, s, x, n, v);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_copy(const double* x, double* y, size_t n, hipStream_t s) {
  TNAD_HIP_TRY(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}
void k_axpby(double a, const double* x, double b, double* y, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(axpby_kernel, dim3(nblk(n)), dim3(NT), 0, s, a, x, b, y, n);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_scale_by_dev(double* x, size_t n, const double* nrm, hipStream_t s) {
  hipLaunchKernelGGL(scale_by_dev_kernel, dim3(nblk(n)), dim3(NT), 0, s, x, n, nrm);
  TNAD_HIP_TRY(hipGetLastError());
}
// out must point at a buffer of at least RED_BLOCKS + 1 doubles; the result lands in out[0].
void k_dot(const double* x, const double* y, size_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(RED_BLOCKS), dim3(NT), 0, s, x, y, n, out + 1);
  hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(NT), 0, s, out + 1, RED_BLOCKS, out);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_sumsq(const double* x, size_t n, double* out, hipStream_t s) { k_dot(x, nullptr, n, out, s); }
void k_edge_sym(const double* T1, int chi, int D, double* T2, hipStream_t s) {
  size_t n = (size_t)chi * D * chi;
  hipLaunchKernelGGL(edge_sym_kernel, dim3(nblk(n)), dim3(NT), 0, s, T1, chi, D, T2);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_col_norms(const double* X, int n, int k, double* nrm, hipStream_t s) {
  hipLaunchKernelGGL(col_norms_kernel, dim3(k), dim3(NT), 0, s, X, n, nrm);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_scale_cols(double* X, int n, int k, const double* scale, int invert, hipStream_t s) {
  hipLaunchKernelGGL(scale_cols_kernel, dim3(nblk((size_t)n * k)), dim3(NT), 0, s, X, n, k, scale, invert);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_normalize_cols(double* X, int n, int k, double* nrm, double rel, unsigned seed, hipStream_t s) {
  hipLaunchKernelGGL(col_norms_kernel, dim3(k), dim3(NT), 0, s, X, n, nrm);
  k_fill_scale_cols(X, n, k, nrm, rel, seed, s);
}
void k_fill_scale_cols(double* X, int n, int k, const double* nrm, double rel, unsigned seed, hipStream_t s) {
  hipLaunchKernelGGL(normalize_fill_cols_kernel, dim3(k), dim3(NT), 0, s, X, n, k, nrm, rel, seed);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_random_fill_flagged(double* X, int n, int k, const int* flag, unsigned seed, hipStream_t s) {
  hipLaunchKernelGGL(random_fill_flagged_kernel, dim3(nblk((size_t)n * k)), dim3(NT), 0, s, X, n, k, flag, seed);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_random_matrix(double* X, size_t n, unsigned seed, hipStream_t s) {
  hipLaunchKernelGGL(random_matrix_kernel, dim3(nblk(n)), dim3(NT), 0, s, X, n, seed);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_init_raw(const double* a, int D, int chi, double* C, double* T, hipStream_t s) {
  size_t n = (size_t)chi * chi + (size_t)chi * D * chi;
  hipLaunchKernelGGL(init_raw_kernel, dim3(nblk(n)), dim3(NT), 0, s, a, D, chi, C, T);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_symmetrize_bulk(const double* x, int d, int ph, double* y, hipStream_t s) {
  int N = d * d * d * d * ph;
  size_t lds = 2 * (size_t)N * sizeof(double);
  if (lds > 64 * 1024) throw Error(TNAD_ERR_SHAPE, "symmetrize: bulk tensor too large");
  hipLaunchKernelGGL(symmetrize_bulk_kernel, dim3(1), dim3(1024), lds, s, x, d, ph, y, 0,
                     (const double*)nullptr);
  TNAD_HIP_TRY(hipGetLastError());
}
// g: cotangent of the symmetrised output; x_orig: the unsymmetrised input
void k_symmetrize_bulk_back2(const double* g, const double* x_orig, int d, int ph, double* out,
                             hipStream_t s) {
  int N = d * d * d * d * ph;
  size_t lds = 2 * (size_t)N * sizeof(double);
  if (lds > 64 * 1024) throw Error(TNAD_ERR_SHAPE, "symmetrize: bulk tensor too large");
  hipLaunchKernelGGL(symmetrize_bulk_kernel, dim3(1), dim3(1024), lds, s, g, d, ph, out, 1, x_orig);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_double_layer(const double* bulk, int d, int ph, double* ap, double* a, hipStream_t s) {
  size_t D4 = (size_t)d * d * d * d * d * d * d * d;
  hipLaunchKernelGGL(double_layer_kernel, dim3(nblk(D4)), dim3(NT), 0, s, bulk, d, ph, ap, a);
  TNAD_HIP_TRY(hipGetLastError());
}
void k_double_layer_back(const double* bulk, int d, int ph, const double* apbar, const double* abar,
                         double* bbar, hipStream_t s) {
  size_t N = (size_t)d * d * d * d * ph;
  hipLaunchKernelGGL(double_layer_back_kernel, dim3(nblk(N, 64)), dim3(64), 0, s, bulk, d, ph, apbar, abar,
                     bbar);
  TNAD_HIP_TRY(hipGetLastError());
}

}  // namespace tnad
