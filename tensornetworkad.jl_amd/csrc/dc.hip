// Dense symmetric eigensolver for m <= 192 (FP64) in four stages, each a HIP kernel:
//
//  * trd_kernel : Householder reduction to tridiagonal form in one workgroup of 32 x 16 threads.  The
//    matrix stays in registers in a cyclic distribution (thread (ty, tx) holds the 6 x 12 elements
//    (ty + 32 a, tx + 16 b)), both triangles, so the product A v of each step is a per-thread partial
//    sum over its twelve columns finished by DPP butterflies across the 16 lanes of the same ty (one
//    row of a wavefront), and the symmetric rank-two update is local.  Per step: p = tau A v through
//    LDS, while the owner lanes of column j + 1 publish it as it stands; after the second barrier every
//    thread updates its elements, and wavefront 0 forms column j + 1 after the update from the published
//    column, v, p and p . v alone, then the next reflector from it, and publishes v and tau (the other
//    wavefronts never repeat its reductions, square root and division); two barriers per step; column
//    blocks that a step cannot touch are not emitted (template block index), finished row blocks are
//    skipped.
//    The reflectors (v with v[j+1] = 1, zeros above) and their tau are written for the back
//    transformation.
//  * dc_kernel : divide and conquer on the tridiagonal matrix T (Cuppen's method), one workgroup.  T is
//    torn recursively into halves down to 1 x 1 leaves, T = diag(T1, T2) + |b| u u^T with
//    u = e_last + sign(b) e_first, and the tree is merged bottom-up, all merges of a level at once:
//    the children's eigenpairs give D + rho z z^T, z = Q^T u / sqrt 2, rho = 2 |b|; entries with
//    rho |z_i| below 8 eps max(|D|, rho |z|) deflate, and so does one of two entries whose Givens
//    rotation (zeroing one z) perturbs the matrix by less than that (Dongarra-Sorensen deflation);
//    the remaining K roots of the secular equation 1 + rho sum z_i^2 / (d_i - lambda) = 0 are found
//    in parallel, each as an offset tau from its nearer pole (so lambda - d_i is formed without
//    cancellation) by the two-pole rational model of the secular function (Li's "middle way"),
//    safeguarded by bisection on a bracket; z is then recomputed from the roots (Gu-Eisenstat, Lowner's
//    formula), which makes the eigenvectors (zhat_i / (d_i - lambda_j))_i orthogonal to working
//    accuracy; the merged eigenvectors are the children's (block-diagonal) eigenvector matrix times
//    that K x K matrix (deflated columns are copied).  The eigenvector matrices of a level live in
//    two m x m ping-pong buffers in global memory (L2-resident), a level's K x K matrices in a third.
//    The top merge stops short of its product: it leaves the children's eigenvectors Z in Qb and its
//    m x m merge matrix M (the K x K matrix scattered to the rows of its columns of Z, a 1 for each
//    deflated column) in U, since H (Z M) = (H Z) M lets the product run over the whole chip.
//  * trd_back_kernel : Z <- H_0 H_1 ... H_(m-3) Z, one wavefront per column of Z (no barriers), four
//    reflectors per wavefront reduction.
//  * dc_final_kernel : W = Z M, a 16 x 16 tile of W per workgroup on the FP64 matrix cores.
// The eigenpairs are then sorted by |lambda| descending with the sign convention of the Jacobi path.
#include "tnad_internal.h"

#include <cfloat>

namespace tnad {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int TD_GX = 16, TD_GY = 32;       // thread grid: 32 rows x 16 columns of threads (512)
constexpr int TD_NT = TD_GX * TD_GY;
static_assert(TD_GX == 8 || TD_GX == 16 || TD_GX == 32, "row sums span half, one or two DPP rows");
constexpr int TD_RB = JACOBI_MAX / TD_GX;   // column elements per thread (12)
constexpr int TD_RA = JACOBI_MAX / TD_GY;   // row elements per thread (4)
static_assert(TD_RB * TD_GX == JACOBI_MAX && TD_RA * TD_GY == JACOBI_MAX, "tridiagonalisation tiling");

// DPP lane move of a double (its two 32-bit halves); lanes outside row_mask keep x.  With every row enabled
// no old value is needed, so the move writes a fresh register (no copy of x ahead of it).
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  int lo, hi;
  if constexpr (ROWMASK == 0xF) {
    lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, ROWMASK, 0xF, false);
    hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
  }
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sums over each row of 16 lanes of N values at once, stage by stage (the lane moves of different
// values interleave); butterflies, so every lane of a row ends with the same bits
template <int N>
__device__ __forceinline__ void row16_sums(double (&x)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) x[a] += dpp_d<0xB1>(x[a]);    // quad_perm (1, 0, 3, 2)
#pragma unroll
  for (int a = 0; a < N; ++a) x[a] += dpp_d<0x4E>(x[a]);    // quad_perm (2, 3, 0, 1)
#pragma unroll
  for (int a = 0; a < N; ++a) x[a] += dpp_d<0x141>(x[a]);   // row_half_mirror
#pragma unroll
  for (int a = 0; a < N; ++a) x[a] += dpp_d<0x140>(x[a]);   // row_mirror
}

// totals of N values over the wavefront, in every lane (the N butterflies interleave)
template <int N>
__device__ __forceinline__ void wave_sums(double (&x)[N]) {
  row16_sums<N>(x);
#pragma unroll
  for (int k = 0; k < N; ++k) x[k] += dpp_d<0x142, 0xA>(x[k]);
#pragma unroll
  for (int k = 0; k < N; ++k) x[k] += dpp_d<0x143, 0x8>(x[k]);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const long long b = __double_as_longlong(x[k]);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    x[k] = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// sum over the wavefront, the same value in every lane
__device__ __forceinline__ double wave_sum(double x) {
  double v[1] = {x};
  row16_sums<1>(v);
  x = v[0];
  x += dpp_d<0x142, 0xA>(x);   // row_bcast15: row 1 += row 0, row 3 += row 2
  x += dpp_d<0x143, 0x8>(x);   // row_bcast31: row 3 += lane 31; lane 63 holds the total
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

struct TrdShared {
  double col[2][JACOBI_MAX];   // column j + 1 as it stands before step j's update (double-buffered by parity)
  double v[2][JACOBI_MAX];     // the reflector v (v[j+1] = 1, zeros elsewhere outside j+1..m-1)
  double p[2][JACOBI_MAX];     // tau A v
  double colb[JACOBI_MAX];     // the same for the last step of a column block (column j + 1 in the next block)
  double tv[2];                // tau
};

// the value of x in lane k (k uniform)
__device__ __forceinline__ double lane_of(double x, int k) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, k), hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// wavefront 0: the reflector H = I - tau v v^T with v[j+1] = 1 that maps column j's rows j+1.. (c: rows lane,
// lane + 64, lane + 128 of the lane) to (beta, 0, ..., 0); publishes v and tau (buffer j & 1) and writes v to
// Vh, tau and beta to the outputs
__device__ __forceinline__ void trd_reflect(const double (&c)[3], int j, int m, int lane, TrdShared& sh,
                                            double* __restrict__ Vh, double* __restrict__ tauo,
                                            double* __restrict__ eo) {
  double sg = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int r = lane + 64 * k;
    sg += (r > j + 1 && r < m) ? c[k] * c[k] : 0.0;
  }
  sg = wave_sum(sg);
  // alpha = c at row j + 1: read from its lane in each of the three slots, then the slot chosen (a select of
  // values, not of an index into c, which would send c to scratch memory)
  const int jr = j + 1, sl = jr >> 6;
  const double a0 = lane_of(c[0], jr & 63), a1 = lane_of(c[1], jr & 63), a2 = lane_of(c[2], jr & 63);
  const double al = sl == 0 ? a0 : (sl == 1 ? a1 : a2);
  double be = al, tv = 0.0, sc = 0.0;
  if (sg > 0.0) {
    const double nr = sqrt(al * al + sg);
    be = al >= 0.0 ? -nr : nr;
    const double d = al - be;             // |d| >= |al|: no cancellation
    const double inv = 1.0 / (be * d);    // tau = (beta - alpha) / beta = -d^2 inv, 1 / d = beta inv
    tv = -d * d * inv;
    sc = be * inv;
  }
  double* vv = sh.v[j & 1];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int r = lane + 64 * k;
    const double v = r == j + 1 ? 1.0 : ((r > j + 1 && r < m) ? c[k] * sc : 0.0);
    vv[r] = v;
    if (r < m) Vh[r + (size_t)j * m] = v;
  }
  if (lane == 0) {
    sh.tv[j & 1] = tv;
    tauo[j] = tv;
    eo[j] = be;
  }
}

// Householder steps j = TD_GX BJ .. min(TD_GX (BJ + 1) - 1, m - 3).  Columns below TD_GX BJ are finished (they lie
// left of every remaining step's column), so only column blocks b >= BJ are touched; row blocks whose
// rows all lie at or above j are skipped: those finished before the block's first step statically (A0),
// the others by a uniform test.  Step j finds v_j and tau_j published (by wavefront 0, during step j - 1, or
// before the first step); every thread forms its share of p = tau A v, and the owners of column j + 1
// publish it as it stands; after the second barrier every thread updates its elements while wavefront 0
// forms column j + 1 after the update from the published column, v, p and (p . v) alone, then the next
// reflector from it, so two barriers per step.
template <int BJ>
__device__ __forceinline__ void trd_block(double (&A)[TD_RA][TD_RB], int m, TrdShared& sh, int tx, int ty,
                                          int lane, int tid, double* __restrict__ Vh, double* __restrict__ tauo,
                                          double* __restrict__ eo) {
  constexpr int A0 = (BJ * TD_GX + 1) / TD_GY;   // row blocks a < A0 end at or above row BJ TD_GX
  const int j0 = BJ * TD_GX, j1 = min(m - 2, j0 + TD_GX);
  for (int j = j0; j < j1; ++j) {
    const int buf = j & 1;
    const double* vv = sh.v[buf];
    double* pv = sh.p[buf];
    const bool next = j + 1 < m - 2;   // uniform: a step j + 1 follows
    __syncthreads();
    const double tv = sh.tv[buf];
    // p = tau A v on the rows below j (finished row blocks are skipped and never stored); column by column,
    // so only the matrix, the row sums and one v entry are live
    double ps[TD_RA] = {};
#pragma unroll
    for (int b = BJ; b < TD_RB; ++b) {
      const double vcb = vv[tx + TD_GX * b];
#pragma unroll
      for (int a = A0; a < TD_RA; ++a) {
        if (TD_GY * a + TD_GY - 1 <= j) continue;   // uniform: every row of the block is finished
        ps[a] = fma(A[a][b], vcb, ps[a]);
      }
    }
    // row sums over the TD_GX lanes of each row: pairs, quads, half rows (8 lanes), rows (16), row pairs (32)
#pragma unroll
    for (int a = A0; a < TD_RA; ++a) ps[a] += dpp_d<0xB1>(ps[a]);
#pragma unroll
    for (int a = A0; a < TD_RA; ++a) ps[a] += dpp_d<0x4E>(ps[a]);
#pragma unroll
    for (int a = A0; a < TD_RA; ++a) ps[a] += dpp_d<0x141>(ps[a]);
    if constexpr (TD_GX >= 16) {
#pragma unroll
      for (int a = A0; a < TD_RA; ++a) ps[a] += dpp_d<0x140>(ps[a]);
    }
    if constexpr (TD_GX == 32) {   // row_bcast15 adds row 0 into row 1, row 2 into row 3
#pragma unroll
      for (int a = A0; a < TD_RA; ++a) ps[a] += dpp_d<0x142, 0xA>(ps[a]);
    }
#pragma unroll
    for (int a = A0; a < TD_RA; ++a) {
      if (TD_GY * a + TD_GY - 1 <= j) continue;
      const int r = ty + TD_GY * a;
      if (tx == (TD_GX == 32 ? 16 : 0)) pv[r] = (r > j && r < m) ? tv * ps[a] : 0.0;   // a lane that holds the row total
    }
    // column j + 1 before this step's update (its rows in finished blocks are never read)
    if (next) {
      double* cn = sh.col[buf];
      if (j + 1 < j0 + TD_GX) {
        if (tx == j + 1 - j0) {
#pragma unroll
          for (int a = A0; a < TD_RA; ++a) cn[ty + TD_GY * a] = A[a][BJ];
        }
      } else if constexpr (BJ + 1 < TD_RB) {   // a buffer of its own: one store merged from the two
        if (tx == 0) {                          // branches would index A dynamically (A in scratch memory)
#pragma unroll
          for (int a = A0; a < TD_RA; ++a) sh.colb[ty + TD_GY * a] = A[a][BJ + 1];
        }
      }
    }
    __syncthreads();
    // w = p - (tau / 2) (p . v) v ;  A -= v w^T + w v^T
    double pw = 0.0;
    for (int r = j + 1 + lane; r < m; r += 64) pw += pv[r] * vv[r];
    pw = wave_sum(pw);
    const double kc = 0.5 * tv * pw;
    if (next && tid < 64) {   // column j + 1 after the update (v[j+1] = 1), then the next reflector
      const double* cn = j + 1 < j0 + TD_GX ? sh.col[buf] : sh.colb;
      const double wj1 = pv[j + 1] - kc;
      double c[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int r = lane + 64 * k;
        const double x = cn[r] - vv[r] * wj1 - (pv[r] - kc * vv[r]);
        c[k] = r >= j + 2 ? x : 0.0;   // rows above hold stale values
      }
      trd_reflect(c, j + 1, m, lane, sh, Vh, tauo, eo);
    }
    // the thread's rows of v and w, then the update column by column (p is written this step only for
    // rows in unskipped blocks: columns at or left of j take w = 0)
    double vr[TD_RA], wr[TD_RA];
#pragma unroll
    for (int a = A0; a < TD_RA; ++a) {
      const int r = ty + TD_GY * a;
      vr[a] = vv[r];
      wr[a] = TD_GY * a + TD_GY - 1 <= j ? 0.0 : pv[r] - kc * vr[a];
    }
#pragma unroll
    for (int b = BJ; b < TD_RB; ++b) {
      const int c = tx + TD_GX * b;
      const double vcb = vv[c], x = pv[c] - kc * vcb, wcb = c > j ? x : 0.0;
#pragma unroll
      for (int a = A0; a < TD_RA; ++a) {
        if (TD_GY * a + TD_GY - 1 <= j) continue;
        A[a][b] = fma(-wr[a], vcb, fma(-vr[a], wcb, A[a][b]));
      }
    }
  }
}

// the column blocks BJ, BJ + 1, ..., TD_RB - 1 in turn
template <int BJ>
__device__ __forceinline__ void trd_blocks(double (&A)[TD_RA][TD_RB], int m, TrdShared& sh, int tx, int ty,
                                           int lane, int tid, double* __restrict__ Vh, double* __restrict__ tauo,
                                           double* __restrict__ eo) {
  trd_block<BJ>(A, m, sh, tx, ty, lane, tid, Vh, tauo, eo);
  if constexpr (BJ + 1 < TD_RB) trd_blocks<BJ + 1>(A, m, sh, tx, ty, lane, tid, Vh, tauo, eo);
}

// H (m x m, ld ldh; the upper triangle is read) -> diagonal dout (m), off-diagonal eout (m - 1),
// reflectors Vh (m x m, column j = v_j) and tauo (m - 2).
__global__ __launch_bounds__(TD_NT) void trd_kernel(const double* __restrict__ H, int ldh, int m,
                                                    double* __restrict__ dout, double* __restrict__ eout,
                                                    double* __restrict__ Vh, double* __restrict__ tauo,
                                                    const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ TrdShared sh;
  const int tid = threadIdx.x, ty = tid / TD_GX, tx = tid % TD_GX, lane = tid & 63;
  double A[TD_RA][TD_RB];
#pragma unroll
  for (int a = 0; a < TD_RA; ++a)
#pragma unroll
    for (int b = 0; b < TD_RB; ++b) {
      const int r = ty + TD_GY * a, c = tx + TD_GX * b;
      const bool in = r < m && c < m;
      const double x = H[in ? min(r, c) + (size_t)max(r, c) * ldh : 0];   // branch-free, clamped address
      A[a][b] = in ? x : 0.0;
    }
  if (m > 2) {   // the first reflector, from column 0 (published, then formed by wavefront 0)
    if (tx == 0) {
#pragma unroll
      for (int a = 0; a < TD_RA; ++a) sh.col[1][ty + TD_GY * a] = A[a][0];
    }
    __syncthreads();
    if (tid < 64) {
      double c[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) c[k] = sh.col[1][lane + 64 * k];
      trd_reflect(c, 0, m, lane, sh, Vh, tauo, eout);
    }
  }
  trd_blocks<0>(A, m, sh, tx, ty, lane, tid, Vh, tauo, eout);
#pragma unroll
  for (int a = 0; a < TD_RA; ++a)
#pragma unroll
    for (int b = 0; b < TD_RB; ++b) {
      const int r = ty + TD_GY * a, c = tx + TD_GX * b;
      if (r == c && r < m) dout[r] = A[a][b];
      if (r == m - 1 && c == m - 2) eout[m - 2] = A[a][b];
    }
}

// ---------------------------------------------------------------------------------------------------
// divide and conquer
// ---------------------------------------------------------------------------------------------------
constexpr int DC_NT = 1024;
constexpr int DC_M = JACOBI_MAX;
constexpr int DC_SEC = 256;   // threads from DC_SEC on work four lanes per root / per entry
constexpr int DC_BIG = 32;    // levels whose widest merge reaches this use the matrix-core product
constexpr int DC_ONE_LANE = 16;   // levels whose merges are at most this wide solve one root per lane

struct DcShared {
  double lam[DC_M];                 // eigenvalue of column g of the current eigenvector matrix
  double ds[DC_M], zs[DC_M];        // merged node: d and z sorted ascending by d
  int src[DC_M];                    // column of the sorted entry
  int nlo[DC_M];                    // lo of the merge node of index g at this level (-1: none)
  double dK[DC_M], zK[DC_M];        // kept entries (slots lo .. lo + K - 1)
  double z2[DC_M];                  // rho zK^2
  int kc[DC_M], ktype[DC_M];        // their columns; 0: top half only, 1: both halves, 2: bottom half only
  double dvl[DC_M];                 // deflated eigenvalues (slots lo + K .. hi - 1)
  int dcol[DC_M];                   // their columns
  int org[DC_M];                    // pole (absolute slot) the root of slot lo + j is measured from
  double tau[DC_M];                 // root = dK[org] + tau
  double zh[DC_M], rn[DC_M];        // recomputed z; 1 / |u_j|
  int upos[DC_M], kcp[DC_M];        // row of kept entry i in the K x K matrix (first its index within
                                    // its type); column of row p
  double rc[DC_M], rs[DC_M];        // deflation rotations of a node (slots lo ..)
  int ra[DC_M], rb[DC_M];
  int nK[DC_M], nT1[DC_M], nT12[DC_M], nRot[DC_M];   // per node at slot lo
  double nRho[DC_M];
  unsigned long long nmd[DC_M], nmz[DC_M];           // per node: bits of max |d|, max |z| (non-negative)
};

// node of index g at depth dd (recursive halving of [0, m), the lower half the smaller); false when it
// is a leaf (a single index, then lo = g, hi = g + 1)
__device__ __forceinline__ bool dc_node(int g, int dd, int m, int& lo, int& hi) {
  lo = 0;
  hi = m;
  for (int l = 0; l < dd; ++l) {
    if (hi - lo <= 1) return false;
    const int mid = lo + ((hi - lo) >> 1);
    if (g < mid) hi = mid;
    else lo = mid;
  }
  return hi - lo >= 2;
}

__device__ __forceinline__ int dc_depth(int m) {
  int d = 0;
  while ((1 << d) < m) ++d;
  return d;
}

// sums / products over the L lanes (L = 1 or 4, aligned) that share a root (butterflies: the same
// bits in all of them)
template <int L>
__device__ __forceinline__ double sumL(double x) {
  if constexpr (L == 4) {
    x += __shfl_xor(x, 1);
    x += __shfl_xor(x, 2);
  } else if constexpr (L == 64) {
    x = wave_sum(x);
  }
  return x;
}

template <int L>
__device__ __forceinline__ double prodL(double x) {
  if constexpr (L == 4) {
    x *= __shfl_xor(x, 1);
    x *= __shfl_xor(x, 2);
  } else if constexpr (L == 64) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) x *= __shfl_xor(x, o);
  }
  return x;
}

// root j (0 <= j < K) of 1 + rho sum_i zk_i^2 / (dk_i - lambda) = 0, dk ascending, four lanes (q = 0..3)
// sharing the sums.  Returns the pole o (index into dk) and lambda - dk[o].
// 1 / x from the hardware estimate and two Newton steps (x finite and nonzero)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}

// root e in (le, he) of c (djl - e)(djr - e) + s1 (djr - e) + s2 (djl - e) = 0 (the secular function
// modelled by its two nearest poles and a constant); false when none lies there
__device__ __forceinline__ bool dc_quad(double c, double djl, double djr, double s1, double s2, double le, double he,
                                        double& e) {
  // (quotients by the refined reciprocal: an iterate needs no correctly rounded division)
  const double qa = c, qb = -(c * (djl + djr) + s1 + s2), qc = c * djl * djr + s1 * djr + s2 * djl;
  if (qa == 0.0) {
    if (qb == 0.0) return false;
    e = -qc * rcp_nr(qb);
    return e > le && e < he;
  }
  const double sq = sqrt(fmax(qb * qb - 4.0 * qa * qc, 0.0));
  const double qq = -0.5 * (qb + copysign(sq, qb));
  e = qq * rcp_nr(qa);
  if (e > le && e < he) return true;
  if (qq == 0.0) return false;
  e = qc * rcp_nr(qq);
  return e > le && e < he;
}

template <int L>
__device__ void dc_secular(const double* dk, const double* z2, int K, int j, int q, int& o, double& t) {
  // fn(i, d_i, rho z_i^2) for every term of this lane
  auto terms = [&](auto&& fn) {
    for (int i = q; i < K; i += L) fn(i, dk[i], z2[i]);
  };
  double lo_t, hi_t;
  if (j < K - 1) {
    const double mdp = 0.5 * (dk[j] + dk[j + 1]);
    double f = 0.0;
    terms([&](int, double d, double w) { f += w * rcp_nr(d - mdp); });
    f = 1.0 + sumL<L>(f);
    o = f > 0.0 ? j : j + 1;
    const double gap = dk[j + 1] - dk[j];
    if (o == j) { lo_t = 0.0; hi_t = 0.5 * gap; }
    else { lo_t = -0.5 * gap; hi_t = 0.0; }
    // first iterate: the two nearest poles exactly, the rest of the sum frozen at the midpoint
    const double djl = dk[j] - dk[o], djr = dk[j + 1] - dk[o];
    const double cc = f - z2[j] / (dk[j] - mdp) - z2[j + 1] / (dk[j + 1] - mdp);
    double e;
    t = dc_quad(cc, djl, djr, z2[j], z2[j + 1], lo_t, hi_t, e) ? e : 0.5 * (lo_t + hi_t);
  } else {
    double zz = 0.0;
    terms([&](int, double, double w) { zz += w; });
    o = j;
    lo_t = 0.0;
    hi_t = sumL<L>(zz);
    t = 0.5 * (lo_t + hi_t);
  }
  const double dorg = dk[o], dj = dk[j] - dorg, dj1 = j < K - 1 ? dk[j + 1] - dorg : 0.0;
  for (int it = 0; it < 80; ++it) {
    // two terms per pass into separate sums, so the two reciprocal chains overlap
    double f = 0.0, af = 0.0, psp = 0.0, php = 0.0, f2 = 0.0, af2 = 0.0, psp2 = 0.0, php2 = 0.0;
    int i = q;
    for (; i + L < K; i += 2 * L) {
      const double r0 = rcp_nr((dk[i] - dorg) - t), r1 = rcp_nr((dk[i + L] - dorg) - t);
      const double w0 = z2[i] * r0, w1 = z2[i + L] * r1;
      f += w0;
      f2 += w1;
      af += fabs(w0);
      af2 += fabs(w1);
      if (i <= j) psp += w0 * r0;
      else php += w0 * r0;
      if (i + L <= j) psp2 += w1 * r1;
      else php2 += w1 * r1;
    }
    if (i < K) {
      const double r = rcp_nr((dk[i] - dorg) - t);
      const double w = z2[i] * r;
      f += w;
      af += fabs(w);
      if (i <= j) psp += w * r;
      else php += w * r;
    }
    f += f2;
    af += af2;
    psp += psp2;
    php += php2;
    f = 1.0 + sumL<L>(f);
    af = sumL<L>(af);
    psp = sumL<L>(psp);
    php = sumL<L>(php);
    if (fabs(f) <= 16.0 * DBL_EPSILON * (1.0 + af) || hi_t - lo_t <= 2.0 * DBL_EPSILON * fmax(fabs(lo_t), fabs(hi_t)))
      break;
    if (f > 0.0) hi_t = t;
    else lo_t = t;
    const double djl = dj - t;
    const double s1 = djl * djl * psp;
    double tn = t;
    bool ok = false;
    if (j < K - 1) {
      const double djr = dj1 - t;
      const double s2 = djr * djr * php;
      const double c = f - s1 * rcp_nr(djl) - s2 * rcp_nr(djr);
      double e;
      if (dc_quad(c, djl, djr, s1, s2, lo_t - t, hi_t - t, e)) { tn = t + e; ok = true; }
    } else {
      const double c = f - s1 / djl;
      if (c != 0.0) {
        const double e = djl + s1 / c;
        if (e > lo_t - t && e < hi_t - t) { tn = t + e; ok = true; }
      }
    }
    if (!ok || !(tn > lo_t && tn < hi_t)) tn = 0.5 * (lo_t + hi_t);
    t = tn;
  }
}

// state of a merge node's deflation scan over its sorted entries: the kept entries so far (K, per type
// c0 top half only, c1 both halves (rotated), c2 bottom half only), the deflated ones (nd), the rotations
// (nr) and the pending entry (the last one not deflated, which a close successor may still rotate away)
struct DcScan {
  int K = 0, nd = 0, nr = 0, c0 = 0, c1 = 0, c2 = 0;
  bool have = false, prot = false;
  double dp = 0.0, zp = 0.0;
  int sp = 0;
};

// deflation threshold of the node at lo: 8 eps max(max |d|, rho max |z|)
__device__ __forceinline__ double dc_tol(const DcShared& sh, int lo, double rho) {
  const double dmx = __longlong_as_double((long long)sh.nmd[lo]), zmx = __longlong_as_double((long long)sh.nmz[lo]);
  return 8.0 * DBL_EPSILON * fmax(dmx, rho * zmx);
}

// keep entry (d, z, column src, rotated?) as kept entry number st.K
__device__ __forceinline__ void dc_keep(DcShared& sh, int lo, int mid, double rho, double d, double z, int src,
                                        bool rot, DcScan& st) {
  const int ty = rot ? 1 : (src < mid ? 0 : 2);
  sh.dK[lo + st.K] = d;
  sh.zK[lo + st.K] = z;
  sh.z2[lo + st.K] = rho * z * z;
  sh.kc[lo + st.K] = src;
  sh.ktype[lo + st.K] = ty;
  sh.upos[lo + st.K] = ty == 0 ? st.c0 : (ty == 1 ? st.c1 : st.c2);
  st.c0 += ty == 0;
  st.c1 += ty == 1;
  st.c2 += ty == 2;
  ++st.K;
}

// sequential scan of the sorted entries from .. hi - 1 (one lane): small z deflates; a close pending /
// new pair is rotated (Givens on the two columns, recorded for the rows of Q) so that the pending entry
// deflates
__device__ __forceinline__ void dc_scan_seq(DcShared& sh, int lo, int hi, int mid, double rho, double tol, int from, DcScan& st) {
  double dnx = sh.ds[from], znx = sh.zs[from];
  int snx = sh.src[from];
  for (int nj = from; nj < hi; ++nj) {
    const double dn = dnx, zn = znx;
    const int sn = snx;
    if (nj + 1 < hi) {   // the next entry's loads are in flight while this one is decided
      dnx = sh.ds[nj + 1];
      znx = sh.zs[nj + 1];
      snx = sh.src[nj + 1];
    }
    if (rho * fabs(zn) <= tol) {
      sh.dvl[hi - 1 - st.nd] = dn;
      sh.dcol[hi - 1 - st.nd] = sn;
      ++st.nd;
      continue;
    }
    if (st.have) {
      const double dt = dn - st.dp;
      if (fabs(dt) <= 2.0 * tol) {   // |dt c s| <= |dt| / 2: only close pairs can deflate
        const double tt = hypot(zn, st.zp);
        const double c = zn / tt, sr = -st.zp / tt;
        if (fabs(dt * c * sr) <= tol) {
          sh.ra[lo + st.nr] = st.sp;
          sh.rb[lo + st.nr] = sn;
          sh.rc[lo + st.nr] = c;
          sh.rs[lo + st.nr] = sr;
          ++st.nr;
          sh.dvl[hi - 1 - st.nd] = st.dp * c * c + dn * sr * sr;
          sh.dcol[hi - 1 - st.nd] = st.sp;
          ++st.nd;
          st.dp = st.dp * sr * sr + dn * c * c;
          st.zp = tt;
          st.sp = sn;
          st.prot = true;
          continue;
        }
      }
      dc_keep(sh, lo, mid, rho, st.dp, st.zp, st.sp, st.prot, st);
    }
    st.have = true;
    st.dp = dn;
    st.zp = zn;
    st.sp = sn;
    st.prot = false;
  }
}

__device__ __forceinline__ void dc_scan_finish(DcShared& sh, int lo, int mid, double rho, DcScan& st) {
  if (st.have) dc_keep(sh, lo, mid, rho, st.dp, st.zp, st.sp, st.prot, st);
  sh.nK[lo] = st.K;
  sh.nT1[lo] = st.c0;
  sh.nT12[lo] = st.c0 + st.c1;
  sh.nRot[lo] = st.nr;
  sh.nRho[lo] = rho;
}

// the scan of the node [lo, hi) by one wavefront: chunks of 64 entries are decided by ballots as long as
// no entry of a chunk lies within 2 tol of its predecessor (then nothing rotates: every entry that is not
// deflated is kept, in order); from the first chunk with such a pair on, lane 0 scans sequentially
// from the state the chunks before it left (the same state the sequential scan would have reached)
__device__ __forceinline__ void dc_scan_wave(DcShared& sh, const double* __restrict__ ev, int lo, int hi, int lane) {
  const int mid = lo + ((hi - lo) >> 1);
  const double rho = 2.0 * fabs(ev[mid - 1]);
  const double tol = dc_tol(sh, lo, rho);
  DcScan st;   // the same in every lane
  const unsigned long long lt = (1ull << lane) - 1ull;
  int h0 = lo;
  for (; h0 < hi; h0 += 64) {
    const int h = h0 + lane;
    const bool in = h < hi;
    const double dn = in ? sh.ds[h] : 0.0, zn = in ? sh.zs[h] : 0.0;
    const int sn = in ? sh.src[h] : 0;
    const bool defl = in && rho * fabs(zn) <= tol;
    const bool kept = in && !defl;
    const unsigned long long mk = __ballot(kept), md = __ballot(defl);
    const unsigned long long below = mk & lt;
    const int pl = below ? 63 - __clzll(below) : 0;
    const double dl = __shfl(dn, pl);
    const bool hasprev = below != 0 || st.have;
    const double dprev = below != 0 ? dl : st.dp;
    if (__ballot(kept && hasprev && fabs(dn - dprev) <= 2.0 * tol)) break;   // uniform
    if (mk) {
      const int last = 63 - __clzll(mk);
      if (st.have) {   // the pending entry of the chunks before: kept
        if (lane == 0) dc_keep(sh, lo, mid, rho, st.dp, st.zp, st.sp, false, st);
        else {
          const int ty = st.sp < mid ? 0 : 2;
          st.c0 += ty == 0;
          st.c2 += ty == 2;
          ++st.K;
        }
      }
      const bool top = sn < mid;
      const unsigned long long mk1 = mk & ~(1ull << last);   // kept now; the last one becomes pending
      const unsigned long long m0 = __ballot(kept && top) & mk1, m2 = __ballot(kept && !top) & mk1;
      if (kept && lane != last) {
        const int K = st.K + __popcll(mk1 & lt);
        sh.dK[lo + K] = dn;
        sh.zK[lo + K] = zn;
        sh.z2[lo + K] = rho * zn * zn;
        sh.kc[lo + K] = sn;
        sh.ktype[lo + K] = top ? 0 : 2;
        sh.upos[lo + K] = top ? st.c0 + __popcll(m0 & lt) : st.c2 + __popcll(m2 & lt);
      }
      st.K += __popcll(mk1);
      st.c0 += __popcll(m0);
      st.c2 += __popcll(m2);
      st.have = true;
      st.prot = false;
      st.dp = __shfl(dn, last);
      st.zp = __shfl(zn, last);
      st.sp = __shfl(sn, last);
    }
    if (defl) {
      const int e = hi - 1 - (st.nd + __popcll(md & lt));
      sh.dvl[e] = dn;
      sh.dcol[e] = sn;
    }
    st.nd += __popcll(md);
  }
  if (lane == 0) {
    if (h0 < hi) dc_scan_seq(sh, lo, hi, mid, rho, tol, h0, st);
    dc_scan_finish(sh, lo, mid, rho, st);
  }
}

// the roots of every merge of a level (threads DC_SEC.., L lanes per root), then z from the roots
// (Lowner's formula) and the order of the kept columns, then the norms of the vectors u_j and the new
// eigenvalues; three barriers
template <int L>
__device__ void dc_roots(DcShared& sh, int m, int tid) {
  const int g = (tid - DC_SEC) / L, q = (tid - DC_SEC) % L;
  const bool mine = tid >= DC_SEC && g < m && sh.nlo[g] >= 0;
  if (mine) {
    const int lo = sh.nlo[g], j = g - lo, K = sh.nK[lo];
    if (j < K) {
      int o;
      double t;
      dc_secular<L>(sh.dK + lo, sh.z2 + lo, K, j, q, o, t);
      if (q == 0) {
        sh.org[lo + j] = lo + o;
        sh.tau[lo + j] = t;
      }
    }
  }
  __syncthreads();
  if (mine) {
    const int lo = sh.nlo[g], i = g - lo, K = sh.nK[lo];
    if (i < K) {
      const double* dk = sh.dK + lo;
      const double di = dk[i], rho = sh.nRho[lo];
      double p = 1.0;
      for (int jj = q; jj < K; jj += L) {
        const double lmd = (sh.dK[sh.org[lo + jj]] - di) + sh.tau[lo + jj];
        const double den = jj == K - 1 ? rho : (jj < i ? dk[jj] - di : dk[jj + 1] - di);
        p *= lmd / den;
      }
      p = prodL<L>(p);
      if (q == 0) {
        sh.zh[lo + i] = copysign(sqrt(fmax(p, 0.0)), sh.zK[lo + i]);
        const int ty = sh.ktype[lo + i];
        const int pos = sh.upos[lo + i] + (ty == 0 ? 0 : (ty == 1 ? sh.nT1[lo] : sh.nT12[lo]));
        sh.upos[lo + i] = pos;
        sh.kcp[lo + pos] = sh.kc[lo + i];
      }
    }
  }
  __syncthreads();
  if (mine) {
    const int lo = sh.nlo[g], j = g - lo, K = sh.nK[lo];
    if (j < K) {
      const double dorg = sh.dK[sh.org[lo + j]], tj = sh.tau[lo + j];
      double s = 0.0;
      for (int i = q; i < K; i += L) {
        const double u = sh.zh[lo + i] / ((sh.dK[lo + i] - dorg) - tj);
        s += u * u;
      }
      s = sumL<L>(s);
      if (q == 0) {
        sh.rn[lo + j] = 1.0 / sqrt(s);
        sh.lam[lo + j] = dorg + tj;
      }
    } else if (q == 0) {
      sh.lam[g] = sh.dvl[g];
    }
  }
  __syncthreads();
}

// the top merge as dc_kernel leaves it (after its deflation and rotations) for dc_sec_kernel, dc_zh_kernel
// and dc_final_kernel: the K kept poles (ascending), their z and rho z^2, their columns of Z, the deflated
// columns (slots K..m-1), and per root its pole, offset and recomputed z
struct DcTop {
  double *dK, *zK, *z2, *tau, *zh;
  int *kc, *dcol, *org, *K;
  double* rho;
};

// one level of merges: Qin holds the children's eigenvectors (block diagonal), Qout receives the
// merged ones; U (m x m) holds each merge's K x K eigenvector matrix of D + rho z z^T, transposed, in
// its diagonal block, rows ordered top-half columns, rotated columns, bottom-half columns.
__device__ void dc_level(DcShared& sh, const double* __restrict__ ev, int m, int dd, double* __restrict__ Qin,
                         double* __restrict__ Qout, double* __restrict__ U, const DcTop& top,
                         double* __restrict__ lam_out, int tid) {
  // children's eigenpairs -> z, sorted d of every merge node
  if (tid < m) {
    const int g = tid;
    int lo, hi;
    if (!dc_node(g, dd, m, lo, hi)) {
      sh.nlo[g] = -1;
      Qout[g + (size_t)g * m] = 1.0;   // a leaf passes through
    } else {
      sh.nlo[g] = lo;
      const int mid = lo + ((hi - lo) >> 1);
      const double be = ev[mid - 1];
      double z = g < mid ? Qin[(mid - 1) + (size_t)g * m] : (be < 0.0 ? -Qin[mid + (size_t)g * m] : Qin[mid + (size_t)g * m]);
      z *= 0.70710678118654752440;
      const double lg = sh.lam[g];
      int rank = 0;
      for (int h = lo; h < hi; ++h) {
        const double lh = sh.lam[h];
        rank += (lh < lg) || (lh == lg && h < g);
      }
      sh.ds[lo + rank] = lg;
      sh.zs[lo + rank] = z;
      sh.src[lo + rank] = g;
      atomicMax(&sh.nmd[lo], (unsigned long long)__double_as_longlong(fabs(lg)));
      atomicMax(&sh.nmz[lo], (unsigned long long)__double_as_longlong(fabs(z)));
    }
  }
  __syncthreads();
  // deflation: a wavefront per merge node while a level has at most 16 of them (ballots decide whole chunks
  // of 64 entries at once until a close pair needs the sequential scan), else one thread per node
  if (dd <= 4) {
    const int wave = tid >> 6, lane = tid & 63;
    if (wave < (1 << dd)) {
      int lo = 0, hi = m;
      for (int l = dd - 1; l >= 0; --l) {
        const int md = lo + ((hi - lo) >> 1);
        if ((wave >> l) & 1) lo = md;
        else hi = md;
      }
      if (hi - lo >= 2) dc_scan_wave(sh, ev, lo, hi, lane);
    }
  } else if (tid < m && sh.nlo[tid] == tid) {
    int lo2, hi;
    dc_node(tid, dd, m, lo2, hi);
    DcScan st;
    const int mid = tid + ((hi - tid) >> 1);
    const double rho = 2.0 * fabs(ev[mid - 1]);
    dc_scan_seq(sh, tid, hi, mid, rho, dc_tol(sh, tid, rho), tid, st);
    dc_scan_finish(sh, tid, mid, rho, st);
  }
  __syncthreads();
  // deflation rotations on the rows of Qin (row-parallel); the secular roots, z from the roots, norms
  if (tid < DC_SEC) {
    const int r = tid;
    if (r < m && sh.nlo[r] >= 0) {
      const int lo = sh.nlo[r], nr = sh.nRot[lo];
      for (int t = 0; t < nr; ++t) {
        const size_t ia = r + (size_t)sh.ra[lo + t] * m, ib = r + (size_t)sh.rb[lo + t] * m;
        const double x = Qin[ia], y = Qin[ib], c = sh.rc[lo + t], s = sh.rs[lo + t];
        Qin[ia] = c * x + s * y;
        Qin[ib] = -s * x + c * y;
      }
    }
  }
  if (dd == 0) {   // the top merge goes on over the whole chip (dc_sec_kernel, dc_zh_kernel, dc_final_kernel)
    const int K = sh.nK[0];
    if (tid < m) {
      if (tid < K) {
        top.dK[tid] = sh.dK[tid];
        top.zK[tid] = sh.zK[tid];
        top.z2[tid] = sh.z2[tid];
        top.kc[tid] = sh.kc[tid];
      } else {
        top.dcol[tid] = sh.dcol[tid];
        lam_out[tid] = sh.dvl[tid];
      }
    }
    if (tid == 0) {
      *top.K = K;
      *top.rho = sh.nRho[0];
    }
    return;
  }
  const int smax = (m + (1 << dd) - 1) >> dd;   // the widest node of this level
  // narrow merges: one lane per root; wide ones: four lanes share each root's sums
  if (smax <= DC_ONE_LANE) dc_roots<1>(sh, m, tid);
  else dc_roots<4>(sh, m, tid);
  // the K x K eigenvector matrices: items (row r, column t) inside the node blocks
  for (int it = tid; it < m * smax; it += DC_NT) {
    const int r = it / smax, t = it % smax;
    const int lo = sh.nlo[r];
    if (lo < 0) continue;
    const int K = sh.nK[lo], i = r - lo;
    if (i >= K || t >= K) continue;
    const double u = sh.zh[lo + i] / ((sh.dK[lo + i] - sh.dK[sh.org[lo + t]]) - sh.tau[lo + t]) * sh.rn[lo + t];
    U[(lo + t) + (size_t)(lo + sh.upos[lo + i]) * m] = u;   // row p = upos of the K x K matrix, stored transposed
  }
  if (tid < m) {
    sh.nmd[tid] = 0;
    sh.nmz[tid] = 0;
  }
  __syncthreads();
  // Qout = Qin[:, kept] U, deflated columns copied.  Levels of narrow merges: one output per item.
  if (smax < DC_BIG) {
    for (int it = tid; it < m * smax; it += DC_NT) {
      const int r = it / smax, t = it % smax;
      const int lo = sh.nlo[r];
      if (lo < 0) continue;
      int lo2, hi;
      dc_node(r, dd, m, lo2, hi);
      if (t >= hi - lo) continue;
      const int K = sh.nK[lo], c = lo + t;
      double acc;
      if (t < K) {
        acc = 0.0;
        for (int p = 0; p < K; ++p) acc += Qin[r + (size_t)sh.kcp[lo + p] * m] * U[c + (size_t)(lo + p) * m];
      } else {
        acc = Qin[r + (size_t)sh.dcol[c] * m];
      }
      Qout[r + (size_t)c * m] = acc;
    }
    __syncthreads();
    return;
  }
  // wide merges: 16 x 16 output tiles on the FP64 matrix cores, one wavefront per tile, operands
  // straight from L2
  const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  int base = 0;
  const int nn = 1 << dd;
  for (int k = 0; k < nn; ++k) {
    int lo = 0, hi = m;
    for (int l = dd - 1; l >= 0; --l) {
      const int mid = lo + ((hi - lo) >> 1);
      if ((k >> l) & 1) lo = mid;
      else hi = mid;
    }
    const int s = hi - lo;
    if (s < 2) continue;   // no merge node here (uniform)
    const int K = sh.nK[lo], mid = lo + (s >> 1);
    const int rt = (s + 15) >> 4, ct = (K + 15) >> 4, ntile = rt * ct;
    for (int tl = wave - base; tl < ntile; tl += DC_NT / 64) {
      if (tl < 0) continue;
      const int r0 = lo + 16 * (tl % rt), t0 = 16 * (tl / rt);
      // rows of the top half meet only top-half and rotated columns, the bottom half the others
      const int p0 = r0 >= mid ? sh.nT1[lo] : 0;
      const int p1 = r0 + 15 < mid ? sh.nT12[lo] : K;
      const int ra = r0 + li;   // A-fragment row of this lane (A[li][lk])
      const int tb = t0 + li;   // B-fragment column of this lane (B[lk][li])
      // clamped addresses, no branches: rows and columns beyond the node only feed outputs that are not
      // stored, and the k-steps beyond p1 are zeroed through b
      const int rac = min(ra, hi - 1), tbc = min(tb, K - 1);
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      // eight k-steps per pass: the column indices first (one LDS wait), then all sixteen operand loads in
      // flight before the first product
      for (int p = p0; p < p1; p += 32) {
        int kk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) kk[u] = sh.kcp[lo + min(p + 4 * u + lk, p1 - 1)];
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          a[u] = Qin[rac + (size_t)kk[u] * m];
          b[u] = U[(lo + tbc) + (size_t)(lo + min(p + 4 * u + lk, p1 - 1)) * m];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], p + 4 * u + lk < p1 ? b[u] : 0.0, acc, 0, 0, 0);
      }
      // output lane layout of the 16 x 16 x 4 FP64 product: register q holds row lk + 4 q, column li
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + lk + 4 * q;
        if (r < hi && tb < K) Qout[r + (size_t)(lo + tb) * m] = acc[q];
      }
    }
    base += ntile;
    const int nd = s - K;
    for (int e = tid; e < nd * s; e += DC_NT) {
      const int r = lo + e % s, t = K + e / s;
      Qout[r + (size_t)(lo + t) * m] = Qin[r + (size_t)sh.dcol[lo + t] * m];
    }
  }
  __syncthreads();
}

// tridiagonal (d, e) -> eigenvalues lam (unsorted); the eigenvectors are Qb M, with the top merge's matrix M
// stored transposed in U (m x m each); Qa: m x m scratch
__global__ __launch_bounds__(DC_NT) void dc_kernel(const double* __restrict__ d, const double* __restrict__ e, int m,
                                                   double* __restrict__ Qa, double* __restrict__ Qb,
                                                   double* __restrict__ U, DcTop top, double* __restrict__ lam_out,
                                                   const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ DcShared sh;
  const int tid = threadIdx.x;
  const int depth = dc_depth(m);
  // the levels alternate between the buffers; the top merge reads Qb (as does dc_final_kernel when m = 1)
  const bool odd = depth > 0 && ((depth - 1) & 1);
  double* Qin = odd ? Qa : Qb;
  double* Qout = odd ? Qb : Qa;
  // Qin = I, Qout = 0; the row and column of each element advance without a division per element
  {
    const int rstep = DC_NT % m, cstep = DC_NT / m;
    int r = tid % m, c = tid / m;
    for (int idx = tid; idx < m * m; idx += DC_NT) {
      Qin[idx] = r == c ? 1.0 : 0.0;
      Qout[idx] = 0.0;
      r += rstep;
      c += cstep;
      if (r >= m) {
        r -= m;
        ++c;
      }
    }
  }
  // leaves: every boundary is a tear, d_i -= |e_(i-1)| + |e_i|
  if (tid < m) {
    double v = d[tid];
    if (tid > 0) v -= fabs(e[tid - 1]);
    if (tid < m - 1) v -= fabs(e[tid]);
    sh.lam[tid] = v;
    sh.nmd[tid] = 0;
    sh.nmz[tid] = 0;
  }
  __syncthreads();
  for (int dd = depth - 1; dd >= 0; --dd) {
    dc_level(sh, e, m, dd, Qin, Qout, U, top, lam_out, tid);
    if (dd == 0) break;
    double* t = Qin;
    Qin = Qout;
    Qout = t;
  }
  if (depth == 0 && tid == 0) {   // m = 1: Qb = 1, one deflated column
    *top.K = 0;
    top.dcol[0] = 0;
    lam_out[0] = sh.lam[0];
  }
}

// the top merge's secular roots, one wavefront per root (all 64 lanes share its sums): root j's pole and
// offset, and its eigenvalue
__global__ __launch_bounds__(64) void dc_sec_kernel(DcTop top, double* __restrict__ lam_out,
                                                    const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int j = blockIdx.x, lane = threadIdx.x;
  const int K = *top.K;
  if (j >= K) return;
  __shared__ double dk[JACOBI_MAX], z2[JACOBI_MAX];
  for (int i = lane; i < K; i += 64) {
    dk[i] = top.dK[i];
    z2[i] = top.z2[i];
  }
  __syncthreads();
  int o;
  double t;
  dc_secular<64>(dk, z2, K, j, lane, o, t);
  if (lane == 0) {
    top.org[j] = o;
    top.tau[j] = t;
    lam_out[j] = dk[o] + t;
  }
}

// the top merge's z recomputed from its roots (Lowner's formula), one wavefront per entry
__global__ __launch_bounds__(64) void dc_zh_kernel(DcTop top, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int i = blockIdx.x, lane = threadIdx.x;
  const int K = *top.K;
  if (i >= K) return;
  const double rho = *top.rho, di = top.dK[i];
  double p = 1.0;
  for (int jj = lane; jj < K; jj += 64) {
    const double lmd = (top.dK[top.org[jj]] - di) + top.tau[jj];
    const double den = jj == K - 1 ? rho : (jj < i ? top.dK[jj] - di : top.dK[jj + 1] - di);
    p *= lmd / den;
  }
  p = prodL<64>(p);
  if (lane == 0) top.zh[i] = copysign(sqrt(fmax(p, 0.0)), top.zK[i]);
}

// W = Z M for the top merge, Z = Qb (the children's eigenvectors, deflation-rotated and back-transformed):
// column t < K of M holds the normalised eigenvector zhat_i / (d_i - lambda_t) of the merge on the rows kc_i,
// a deflated column t >= K is column dcol_t of Z.  A 16 x 16 tile of W per workgroup on the FP64 matrix
// cores, over the whole chip; each workgroup forms the norms of its 16 columns and the entries of M it needs.
__global__ __launch_bounds__(64) void dc_final_kernel(const double* __restrict__ Z, DcTop top, int m,
                                                      double* __restrict__ W, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double dk[JACOBI_MAX], zh[JACOBI_MAX];
  __shared__ int kc[JACOBI_MAX];
  const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
  const int r0 = 16 * blockIdx.x, t0 = 16 * blockIdx.y;
  const int K = *top.K;
  for (int i = lane; i < K; i += 64) {
    dk[i] = top.dK[i];
    zh[i] = top.zh[i];
    kc[i] = top.kc[i];
  }
  __syncthreads();
  // this lane's column t = t0 + li: its root's pole and offset, and its norm (four lanes per column)
  const int tb = t0 + li;
  const bool kept = tb < K;
  const double dorg = kept ? dk[top.org[tb]] : 0.0, tt = kept ? top.tau[tb] : 0.0;
  double sq = 0.0;
  if (kept)
    for (int i = lk; i < K; i += 4) {
      const double u = zh[i] / ((dk[i] - dorg) - tt);
      sq += u * u;
    }
  sq += __shfl_xor(sq, 16);
  sq += __shfl_xor(sq, 32);
  const double rn = kept ? 1.0 / sqrt(sq) : 0.0;
  const int ra = r0 + li;
  const int rac = min(ra, m - 1);   // rows beyond m only feed outputs that are not stored
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int p = 0; p < K; p += 64) {
    double a[16], b[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int i = min(p + 4 * u + lk, K - 1);
      a[u] = Z[rac + (size_t)kc[i] * m];
      b[u] = zh[i] / ((dk[i] - dorg) - tt) * rn;
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 16; ++u)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], (kept && p + 4 * u + lk < K) ? b[u] : 0.0, acc, 0, 0, 0);
  }
  // register q holds row lk + 4 q, column li; deflated columns are copied
  const int src = kept ? 0 : (tb < m ? top.dcol[tb] : 0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = r0 + lk + 4 * q;
    if (r < m && tb < m) W[r + (size_t)tb * m] = kept ? acc[q] : Z[r + (size_t)src * m];
  }
}


// Z <- H_0 H_1 ... H_(m-3) Z in place, one wavefront per column.  The reflectors are applied four at a
// time: with s_q = tau_q (u_q . z - sum_(p<q) s_p u_q . u_p) for the block's reflectors u_0..u_3 in the
// order of application, z - sum_q s_q u_q is the product of the four; the four products with z share one
// reduction, and the six among the u come from trd_gram_kernel (lanes 0..5 of one register per block).
// Each lane needs only its own entries of the reflectors (rows lane, lane + 64, lane + 128), so they are
// prefetched in chunks of 16 reflectors into registers, one chunk ahead (two register sets, the loop
// unrolled by two); the reflector count is padded to whole pairs of chunks with tau = 0.
constexpr int BK_Q = 16;

// the six products u_q . u_p (p < q) among each block of four reflectors, in the order the back-transformation
// uses them (u1.u0, u2.u0, u2.u1, u3.u0, u3.u1, u3.u2), once for all columns: one wavefront per block, over
// the padded reflector count (a padding reflector is column 0 again, finite, with tau = 0)
__global__ __launch_bounds__(64) void trd_gram_kernel(const double* __restrict__ Vh, int m, double* __restrict__ gram,
                                                      const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int b = blockIdx.x, lane = threadIdx.x;
  double u[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = m - 3 - (4 * b + q);
    const double* v = Vh + (size_t)(j >= 0 ? j : 0) * m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int r = lane + 64 * k;
      const double x = v[r < m ? r : m - 1];
      u[q][k] = r < m ? x : 0.0;
    }
  }
  double d[6];
  int n = 0;
#pragma unroll
  for (int q = 1; q < 4; ++q)
#pragma unroll
    for (int p = 0; p < q; ++p, ++n) d[n] = u[q][0] * u[p][0] + u[q][1] * u[p][1] + u[q][2] * u[p][2];
  wave_sums<6>(d);
  if (lane < 6) {
    double x = d[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) x = lane == k ? d[k] : x;
    gram[6 * b + lane] = x;
  }
}

__device__ __forceinline__ void back_load(const double* __restrict__ Vh, const double* __restrict__ tau,
                                          const double* __restrict__ gram, int m, int q0, int lane,
                                          double (&u)[BK_Q][3], double (&t)[BK_Q], double (&g)[BK_Q / 4]) {
  const int nblk = ((m - 2 + 2 * BK_Q - 1) / (2 * BK_Q)) * (BK_Q / 2);   // blocks trd_gram_kernel wrote
#pragma unroll
  for (int b = 0; b < BK_Q / 4; ++b)   // lane k: product k (a prefetch past the last block rereads it, unused)
    g[b] = gram[6 * min(q0 / 4 + b, nblk - 1) + (lane < 6 ? lane : 5)];
  // raw loads from clamped addresses (no select on a loaded value here, which would wait for it): a padding
  // reflector gets tau = 0, so whatever it holds has weight 0, and rows >= m are masked where they are used
#pragma unroll
  for (int i = 0; i < BK_Q; ++i) {
    const int j = m - 3 - (q0 + i);   // reflectors in the order of application
    const bool ok = j >= 0;
    const double tj = tau[ok ? j : 0];
    t[i] = ok ? tj : 0.0;
    const double* v = Vh + (size_t)(ok ? j : 0) * m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int r = lane + 64 * k;
      u[i][k] = v[r < m ? r : m - 1];
    }
  }
}

__device__ __forceinline__ void back_apply(double (&z)[3], const double (&uraw)[BK_Q][3], const double (&t)[BK_Q],
                                           const double (&g)[BK_Q / 4], const bool (&rk)[3]) {
#pragma unroll
  for (int b = 0; b < BK_Q; b += 4) {
    double u[BK_Q][3];
#pragma unroll
    for (int q = b; q < b + 4; ++q)
#pragma unroll
      for (int k = 0; k < 3; ++k) u[q][k] = rk[k] ? uraw[q][k] : 0.0;
    double d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = u[b + q][0] * z[0] + u[b + q][1] * z[1] + u[b + q][2] * z[2];
    wave_sums<4>(d);
    const double gb = g[b / 4];
    const double g10 = lane_of(gb, 0), g20 = lane_of(gb, 1), g21 = lane_of(gb, 2), g30 = lane_of(gb, 3),
                 g31 = lane_of(gb, 4), g32 = lane_of(gb, 5);
    const double s0 = t[b] * d[0];
    const double s1 = t[b + 1] * (d[1] - s0 * g10);
    const double s2 = t[b + 2] * (d[2] - s0 * g20 - s1 * g21);
    const double s3 = t[b + 3] * (d[3] - s0 * g30 - s1 * g31 - s2 * g32);
#pragma unroll
    for (int k = 0; k < 3; ++k) z[k] -= s0 * u[b][k] + s1 * u[b + 1][k] + s2 * u[b + 2][k] + s3 * u[b + 3][k];
  }
}

// one wavefront per SIMD is all the grid has, so the scheduler may spend every register on the prefetch
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void trd_back_kernel(
    const double* __restrict__ Vh, const double* __restrict__ tau, const double* __restrict__ gram, int m,
    double* __restrict__ Z, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int c = blockIdx.x, lane = threadIdx.x;
  double* zc = Z + (size_t)c * m;
  double z[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) z[k] = lane + 64 * k < m ? zc[lane + 64 * k] : 0.0;
  const int nr = m - 2;
  const bool rk[3] = {lane < m, lane + 64 < m, lane + 128 < m};
  double ua[BK_Q][3], ta[BK_Q], ga[BK_Q / 4], ub[BK_Q][3], tb[BK_Q], gb[BK_Q / 4];
  back_load(Vh, tau, gram, m, 0, lane, ua, ta, ga);
  // whole pairs of chunks, no exit between a chunk's loads and their use (a conditional one would let the
  // compiler sink the loads past the other chunk's work)
  // (the scheduling barriers keep the machine scheduler from moving the loads down to their uses)
  for (int q = 0; q < nr; q += 2 * BK_Q) {
    back_load(Vh, tau, gram, m, q + BK_Q, lane, ub, tb, gb);
    __builtin_amdgcn_sched_barrier(0);
    back_apply(z, ua, ta, ga, rk);
    back_load(Vh, tau, gram, m, q + 2 * BK_Q, lane, ua, ta, ga);
    __builtin_amdgcn_sched_barrier(0);
    back_apply(z, ub, tb, gb, rk);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (lane + 64 * k < m) zc[lane + 64 * k] = z[k];
}

// the work buffer: Vh, Qa, Qb, U (m x m each), d, e, tau, lambda (m each), the reflector block products
struct DcWork {
  double *Vh, *Qa, *Qb, *U, *dv, *ev, *tau, *lam, *gram;
  DcTop top;
};

DcWork dc_work(double* w, int m) {
  const size_t mm = (size_t)m * m;
  DcWork d;
  d.Vh = w;
  d.Qa = d.Vh + mm;
  d.Qb = d.Qa + mm;
  d.U = d.Qb + mm;
  d.dv = d.U + mm;
  d.ev = d.dv + m;
  d.tau = d.ev + m;
  d.lam = d.tau + m;
  d.gram = d.lam + m;   // 6 per block of four reflectors, the count padded to 32: at most 1.5 m + 44
  double* t = d.gram + 2 * (size_t)m + 48;   // the top merge: 5 m doubles, rho, then 3 m + 1 ints
  d.top.dK = t;
  d.top.zK = t + m;
  d.top.z2 = t + 2 * m;
  d.top.tau = t + 3 * m;
  d.top.zh = t + 4 * m;
  d.top.rho = t + 5 * m;
  int* ti = reinterpret_cast<int*>(t + 5 * m + 1);
  d.top.kc = ti;
  d.top.dcol = ti + m;
  d.top.org = ti + 2 * m;
  d.top.K = ti + 3 * m;
  return d;
}

// blocks of four reflectors that trd_back_kernel reads (m - 2 padded to a multiple of 32)
int dc_gram_blocks(int m) { return m > 2 ? ((m - 2 + 2 * BK_Q - 1) / (2 * BK_Q)) * (BK_Q / 2) : 0; }

// the divide and conquer: the merge tree below the top merge in one workgroup, the top merge's roots and z
// over the whole chip
void dc_solve(const DcWork& w, int m, const int* gate, hipStream_t s) {
  hipLaunchKernelGGL(dc_kernel, dim3(1), dim3(DC_NT), 0, s, w.dv, w.ev, m, w.Qa, w.Qb, w.U, w.top, w.lam, gate);
  TNAD_HIP_TRY(hipGetLastError());
  if (m > 1) {
    hipLaunchKernelGGL(dc_sec_kernel, dim3(m), dim3(64), 0, s, w.top, w.lam, gate);
    TNAD_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(dc_zh_kernel, dim3(m), dim3(64), 0, s, w.top, gate);
    TNAD_HIP_TRY(hipGetLastError());
  }
}

// the eigenvectors: block products and back-transformation of the children's eigenvectors, then the top
// merge's product (into Qa)
void dc_vectors(const DcWork& w, int m, const int* gate, hipStream_t s) {
  if (m > 2) {
    hipLaunchKernelGGL(trd_gram_kernel, dim3(dc_gram_blocks(m)), dim3(64), 0, s, w.Vh, m, w.gram, gate);
    TNAD_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(trd_back_kernel, dim3(m), dim3(64), 0, s, w.Vh, w.tau, w.gram, m, w.Qb, gate);
    TNAD_HIP_TRY(hipGetLastError());
  }
  const int nt = (m + 15) / 16;
  hipLaunchKernelGGL(dc_final_kernel, dim3(nt, nt), dim3(64), 0, s, w.Qb, w.top, m, w.Qa, gate);
  TNAD_HIP_TRY(hipGetLastError());
}

}  // namespace

size_t syev_dc_work_doubles(int m) { return 4 * (size_t)m * m + 13 * (size_t)m + 56; }

void syev_dc(const double* H, int ldh, int m, double* evals, double* V, double* work, hipStream_t s,
             const int* gate) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "syev_dc: size out of range");
  const DcWork w = dc_work(work, m);
  hipLaunchKernelGGL(trd_kernel, dim3(1), dim3(TD_NT), 0, s, H, ldh, m, w.dv, w.ev, w.Vh, w.tau, gate);
  TNAD_HIP_TRY(hipGetLastError());
  dc_solve(w, m, gate, s);
  dc_vectors(w, m, gate, s);
  sort_eigenpairs(w.lam, w.Qa, m, evals, V, gate, s);
}

void syev_dc_time(const double* H, int m, int reps, double* work, hipStream_t s, float ms[3]) {
  if (m < 3 || m > JACOBI_MAX || reps < 1) throw Error(TNAD_ERR_ARG, "syev_dc_time: bad sizes");
  const DcWork w = dc_work(work, m);
  double* V = work + syev_dc_work_doubles(m);   // m x m sorted eigenvectors
  double* evals = V + (size_t)m * m;
  hipEvent_t e[4];
  for (auto& x : e) TNAD_HIP_TRY(hipEventCreate(&x));
  float acc[3] = {0.f, 0.f, 0.f};
  for (int r = 0; r <= reps; ++r) {   // the first run warms up
    TNAD_HIP_TRY(hipEventRecord(e[0], s));
    hipLaunchKernelGGL(trd_kernel, dim3(1), dim3(TD_NT), 0, s, H, m, m, w.dv, w.ev, w.Vh, w.tau, (const int*)nullptr);
    TNAD_HIP_TRY(hipEventRecord(e[1], s));
    dc_solve(w, m, nullptr, s);
    TNAD_HIP_TRY(hipEventRecord(e[2], s));
    dc_vectors(w, m, nullptr, s);
    sort_eigenpairs(w.lam, w.Qa, m, evals, V, nullptr, s);
    TNAD_HIP_TRY(hipEventRecord(e[3], s));
    TNAD_HIP_TRY(hipEventSynchronize(e[3]));
    if (r == 0) continue;
    for (int k = 0; k < 3; ++k) {
      float t = 0.f;
      TNAD_HIP_TRY(hipEventElapsedTime(&t, e[k], e[k + 1]));
      acc[k] += t;
    }
  }
  for (auto& x : e) (void)hipEventDestroy(x);
  for (int k = 0; k < 3; ++k) ms[k] = acc[k] / reps;
}

}  // namespace tnad
