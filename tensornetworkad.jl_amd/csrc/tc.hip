// Generic strided tensor contraction on CDNA4 matrix cores.
//
//   C[m,n] = alpha * sum_k A[m,k] * B[k,n] + beta * C[m,n]
//
// with every index split into two sub-indices carrying arbitrary element
// strides (TcDesc).  This lets each einsum of the CTMRG step, its adjoint and
// the energy chain run as one launch straight from the tensors' natural
// column-major layout: the permutations live in the tile loaders' address
// arithmetic instead of separate transpose passes.
//
// Tiling: 64x64 output tile per 256-thread workgroup (4 waves in 2x2), each
// wave owns a 32x32 sub-tile made of 2x2 MFMA 16x16 blocks.  Operands are
// staged through LDS with a register prefetch of the next K-tile.
// Split-K: a contraction with few output tiles and a long K (the Gram and
// projection products of the eigensolver, the bulk cotangent of the adjoint)
// would leave most of the 256 CUs idle, so tc_launch cuts K into slices along
// grid z; every slice writes its partial tile to a workspace and a second
// kernel sums the slices in a fixed order (deterministic) and applies
// alpha / beta to the strided output.
//   F64 : v_mfma_f64_16x16x4f64       (BK = 16)
//   F32 : v_mfma_f32_16x16x4f32       (BK = 16)
//   BF16: v_mfma_f32_16x16x16bf16_1k  (BK = 32, fp32 in HBM, bf16 in LDS,
//                                      fp32 accumulation)
#include <algorithm>
#include <cstdint>
#include <cstdlib>

#include "tnad_internal.h"

namespace tnad {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

TcDesc gemm_desc(char ta, char tb, int M, int N, int K, long long lda, long long ldb, long long ldc,
                 double alpha, double beta) {
  TcDesc d;
  d.M0 = M; d.N0 = N; d.K0 = K;
  if (ta == 'N') { d.am0 = 1; d.ak0 = lda; } else { d.am0 = lda; d.ak0 = 1; }
  if (tb == 'N') { d.bk0 = 1; d.bn0 = ldb; } else { d.bk0 = ldb; d.bn0 = 1; }
  d.cm0 = 1; d.cn0 = ldc;
  d.alpha = alpha; d.beta = beta;
  return d;
}

namespace {

constexpr int BM = 64, BN = 64, NT = 256;
// K-tile of the FP32 / FP64 kernel: 16 for grids that fill the chip (occupancy matters there), 32 for
// grids of fewer than 256 workgroups, which wait on global-load latency once per K-tile
constexpr int TC_BK = 16, TC_BK_DEEP = 32, TC_DEEP_BELOW = 256;
constexpr int TC_BIG_MIN_WG = 1024;  // 128 x 128 tiles only at 4+ workgroups per CU (slower at 1 per CU)
constexpr int TC_SMALL_BELOW = 128;  // 32 x 32 tiles below this many 64 x 64 workgroups

__device__ __forceinline__ long long off2(int i, int I0, long long s0, long long s1) {
  int i0 = i % I0;
  int i1 = i / I0;
  return (long long)i0 * s0 + (long long)i1 * s1;
}

// Position of one loader along the contraction index k = k0 + K0 * k1: (k0, offset), advanced by a
// K-tile per step without dividing (k0 wraps, carrying into k1 via delta = s1 - K0 * s0).
struct KPos {
  int k0;
  long long off;
  __device__ __forceinline__ void init(int k, int K0, long long s0, long long s1) {
    k0 = k % K0;
    off = (long long)k0 * s0 + (long long)(k / K0) * s1;
  }
  __device__ __forceinline__ void advance(int step, int K0, long long s0, long long delta) {
    k0 += step;
    off += (long long)step * s0;
    while (k0 >= K0) {
      k0 -= K0;
      off += delta;
    }
  }
};

template <typename T> struct MfmaOp;
// Accumulator lane layout of the 16x16 output block (lane l, accumulator register r):
//   f64 16x16x4 : row = (l / 16) + 4 r, column = l % 16
//   f32 16x16x4 : row = 4 (l / 16) + r, column = l % 16
// (operands for both: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]).
template <> struct MfmaOp<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct MfmaOp<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ---------------------------------------------------------------------------
// F64 / F32 kernel (k-major LDS tiles, one element per lane per MFMA).  TILE x TILE output tile per
// workgroup of 4 waves in 2 x 2; each wave owns a (TILE/2)^2 block of (TILE/32)^2 MFMA 16x16 blocks,
// so TILE = 128 does twice the MFMA work per LDS fragment read of TILE = 64.
// ---------------------------------------------------------------------------
// PF: K-tiles in flight in registers (1: the loads of tile t + 2 are issued while tile t is multiplied and
// stored when tile t + 1 is; 2: those of tile t + 3, two register sets alternating, for operands whose
// loads come from beyond the XCD's L2 and need more than one K-tile of cover).
template <typename T, int BK, int TILE, int PF = 1>
__global__ __launch_bounds__(NT, PF == 2 ? (sizeof(T) == 8 ? 3 : 4) : 1) void tc_kernel(TcDesc d, const T* __restrict__ A,
                                                const T* __restrict__ B, T* __restrict__ C,
                                                T* __restrict__ Wsp) {
  constexpr int PAD = 16;
  constexpr int NR = TILE * BK / NT;   // loads per thread and operand per K-tile
  constexpr int FB = TILE / 32;        // MFMA blocks per wave and dimension
  static_assert(NR * NT == TILE * BK && NT % TILE == 0 && NT % BK == 0, "loader coverage");
  __shared__ T As[2][BK][TILE + PAD];   // double-buffered: K-tile t in buffer t % 2
  __shared__ T Bs[2][BK][TILE + PAD];

  const int M = d.M0 * d.M1, N = d.N0 * d.N1;
  const int tile_m = blockIdx.x * TILE, tile_n = blockIdx.y * TILE;
  // grid z: batch index, or split-K slice (batch 1) when a workspace is given
  const int bz = Wsp ? 0 : blockIdx.z;
  const int kbeg = Wsp ? blockIdx.z * d.kchunk : 0;
  const int K = Wsp ? min(d.K0 * d.K1, kbeg + d.kchunk) : d.K0 * d.K1;
  A += (long long)bz * d.abat;
  B += (long long)bz * d.bbat;
  C += (long long)bz * d.cbat;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // A loader mapping: consecutive threads along m (m contiguous in memory) or along k
  const bool a_mfast = (d.am0 == 1 && d.M0 > 1) || d.ak0 != 1;
  int a_mi[NR], a_ki[NR];
  long long a_moff[NR];
  bool a_mok[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (a_mfast) { a_mi[r] = t % TILE; a_ki[r] = t / TILE + (NT / TILE) * r; }
    else         { a_ki[r] = t % BK; a_mi[r] = t / BK + (NT / BK) * r; }
    int m = tile_m + a_mi[r];
    a_mok[r] = m < M;
    a_moff[r] = a_mok[r] ? off2(m, d.M0, d.am0, d.am1) : 0;
  }
  // B loader mapping
  const bool b_nfast = (d.bn0 == 1 && d.N0 > 1) || d.bk0 != 1;
  int b_ni[NR], b_ki[NR];
  long long b_noff[NR];
  bool b_nok[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (b_nfast) { b_ni[r] = t % TILE; b_ki[r] = t / TILE + (NT / TILE) * r; }
    else         { b_ki[r] = t % BK; b_ni[r] = t / BK + (NT / BK) * r; }
    int n = tile_n + b_ni[r];
    b_nok[r] = n < N;
    b_noff[r] = b_nok[r] ? off2(n, d.N0, d.bn0, d.bn1) : 0;
  }

  T ra0[NR], rb0[NR], ra1[NR], rb1[NR];
  KPos a_kp[NR], b_kp[NR];
  const long long a_dk = d.ak1 - (long long)d.K0 * d.ak0, b_dk = d.bk1 - (long long)d.K0 * d.bk0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    a_kp[r].init(kbeg + a_ki[r], d.K0, d.ak0, d.ak1);
    b_kp[r].init(kbeg + b_ki[r], d.K0, d.bk0, d.bk1);
  }
  // loads the K-tile starting at k0 into (xa, xb), then moves the loaders one tile on (tiles are
  // loaded in order)
  auto gload = [&](T (&xa)[NR], T (&xb)[NR], int k0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      int k = k0 + a_ki[r];
      T v = T(0);
      if (a_mok[r] && k < K) v = A[a_moff[r] + a_kp[r].off];
      xa[r] = v;
      int kb = k0 + b_ki[r];
      T w = T(0);
      if (b_nok[r] && kb < K) w = B[b_noff[r] + b_kp[r].off];
      xb[r] = w;
      a_kp[r].advance(BK, d.K0, d.ak0, a_dk);
      b_kp[r].advance(BK, d.K0, d.bk0, b_dk);
    }
  };

  typedef typename MfmaOp<T>::acc_t acc_t;
  acc_t acc[FB][FB];
#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // Software pipeline over K-tiles with two LDS buffers: while tile t is multiplied out of buffer t % 2,
  // tile t + 1 (loaded into registers earlier) is stored into the other buffer and the loads of tile
  // t + 1 + PF are issued.  One barrier per K-tile: it publishes tile t + 1 and retires the reads of
  // tile t, whose buffer the next step refills.
  auto stage = [&](const T (&xa)[NR], const T (&xb)[NR], int buf) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      As[buf][a_ki[r]][a_mi[r]] = xa[r];
      Bs[buf][b_ki[r]][b_ni[r]] = xb[r];
    }
  };
  auto mfma_tile = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[FB], bf[FB];
#pragma unroll
      for (int i = 0; i < FB; ++i) af[i] = As[buf][kk + (lane >> 4)][wm * (TILE / 2) + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FB; ++j) bf[j] = Bs[buf][kk + (lane >> 4)][wn * (TILE / 2) + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FB; ++i)
#pragma unroll
        for (int j = 0; j < FB; ++j) acc[i][j] = MfmaOp<T>::run(af[i], bf[j], acc[i][j]);
    }
  };
  gload(ra0, rb0, kbeg);
  stage(ra0, rb0, 0);
  if (kbeg + BK < K) gload(ra1, rb1, kbeg + BK);
  if (PF == 2 && kbeg + 2 * BK < K) gload(ra0, rb0, kbeg + 2 * BK);
  __syncthreads();
  if constexpr (PF == 1) {
    int buf = 0;
    for (int k0 = kbeg; k0 < K; k0 += BK, buf ^= 1) {
      mfma_tile(buf);
      if (k0 + BK < K) {
        stage(ra1, rb1, buf ^ 1);
        if (k0 + 2 * BK < K) gload(ra1, rb1, k0 + 2 * BK);
      }
      __syncthreads();
    }
  } else {
    // unrolled by two, so the LDS buffer and the register set of every step are fixed: an even tile is
    // multiplied out of buffer 0 while tile t + 1 comes from set 1, an odd one out of buffer 1 while
    // tile t + 1 comes from set 0
    for (int k0 = kbeg; k0 < K;) {
      mfma_tile(0);
      if (k0 + BK < K) {
        stage(ra1, rb1, 1);
        if (k0 + 3 * BK < K) gload(ra1, rb1, k0 + 3 * BK);
      }
      __syncthreads();
      k0 += BK;
      if (k0 >= K) break;
      mfma_tile(1);
      if (k0 + BK < K) {
        stage(ra0, rb0, 0);
        if (k0 + 3 * BK < K) gload(ra0, rb0, k0 + 3 * BK);
      }
      __syncthreads();
      k0 += BK;
    }
  }

  // epilogue: column lane%16 of each 16x16 block, rows per MfmaOp<T>::row
  const T alpha = (T)d.alpha, beta = (T)d.beta;
  T* Wz = Wsp ? Wsp + (size_t)blockIdx.z * M * N : nullptr;   // this slice's M x N partial
#pragma unroll
  for (int j = 0; j < FB; ++j) {
    int n = tile_n + wn * (TILE / 2) + j * 16 + (lane & 15);
    if (n >= N) continue;
    long long noff = off2(n, d.N0, d.cn0, d.cn1);
#pragma unroll
    for (int i = 0; i < FB; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = tile_m + wm * (TILE / 2) + i * 16 + MfmaOp<T>::row(lane, r);
        if (m >= M) continue;
        if (Wz) {
          Wz[m + (size_t)M * n] = (T)acc[i][j][r];
          continue;
        }
        long long o = noff + off2(m, d.M0, d.cm0, d.cm1);
        T v = alpha * (T)acc[i][j][r];
        if (d.beta != 0.0) v += beta * C[o];
        C[o] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// F64 / F32 kernel with 16-byte operand loads (launch_vec decides when it applies): each operand is read
// in vectors of V = 16 / sizeof(T) elements along its unit-stride index (m for A or n for B when that
// index has unit stride, else k), so a thread issues V times fewer loads and no per-element address
// arithmetic; the K-tile never straddles a K0 run (K0 a multiple of BK, or a single K sub-index) and K
// is a multiple of BK, so the K loop carries no bounds tests.  TILE x TILE output tile (64, K-tile 16;
// or 32, K-tile 32, for grids too small to fill the chip with 64 x 64 tiles), 4 waves in 2 x 2,
// double-buffered LDS (one barrier per K-tile), the next K-tile's vectors in flight in registers while
// the current one is multiplied.
// ---------------------------------------------------------------------------
// 16-byte operand vectors as clang vector types (kept in registers; element e is a compile-time
// constant after unrolling)
typedef double d2 __attribute__((ext_vector_type(2)));
template <typename VT> __device__ __forceinline__ auto vec_elem(const VT& v, int e) { return v[e]; }

template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef d2 type; static constexpr int V = 2; };
template <> struct Vec16<float> { typedef f4 type; static constexpr int V = 4; };

template <typename T, int TILE, int BK, bool A_MV, bool B_NV>
__global__ __launch_bounds__(NT, TILE == 128 ? 1 : (sizeof(T) == 8 ? 3 : 4)) void tc_kernel_vec(TcDesc d, const T* __restrict__ A,
                                                                         const T* __restrict__ B,
                                                                         T* __restrict__ C, T* __restrict__ Wsp) {
  // row pitch of the LDS tiles: 16-byte rows; FP32: the four k-rows a fragment read touches start 16
  // banks apart, so its 64 lanes meet no bank twice
  constexpr int PAD = sizeof(T) == 8 ? 4 : 16;
  typedef typename Vec16<T>::type VT;
  constexpr int V = Vec16<T>::V;
  constexpr int NVEC = TILE * BK / (NT * V);   // vectors per thread and operand per K-tile
  constexpr int FB = TILE / 32;
  static_assert(NVEC >= 1 && NVEC * NT * V == TILE * BK, "loader coverage");
  __shared__ __attribute__((aligned(16))) T As[2][BK][TILE + PAD];
  __shared__ __attribute__((aligned(16))) T Bs[2][BK][TILE + PAD];

  const int M = d.M0 * d.M1, N = d.N0 * d.N1;
  const int tile_m = blockIdx.x * TILE, tile_n = blockIdx.y * TILE;
  const int bz = Wsp ? 0 : blockIdx.z;
  const int kbeg = Wsp ? blockIdx.z * d.kchunk : 0;
  const int K = Wsp ? min(d.K0 * d.K1, kbeg + d.kchunk) : d.K0 * d.K1;
  A += (long long)bz * d.abat;
  B += (long long)bz * d.bbat;
  C += (long long)bz * d.cbat;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // loader slots: vectors along m (n) cover a K-tile row by TILE / V threads; vectors along k cover
  // an m (n) column by BK / V threads
  int a_m[NVEC], a_k[NVEC], b_n[NVEC], b_k[NVEC];
  long long a_off[NVEC], b_off[NVEC];
  bool a_ok[NVEC], b_ok[NVEC];
  KPos a_kp[NVEC], b_kp[NVEC];
  const long long a_dk = d.ak1 - (long long)d.K0 * d.ak0, b_dk = d.bk1 - (long long)d.K0 * d.bk0;
#pragma unroll
  for (int r = 0; r < NVEC; ++r) {
    if (A_MV) { a_m[r] = (t % (TILE / V)) * V; a_k[r] = t / (TILE / V) + r * (NT * V / TILE); }
    else      { a_k[r] = (t % (BK / V)) * V;   a_m[r] = t / (BK / V) + r * (NT * V / BK); }
    if (B_NV) { b_n[r] = (t % (TILE / V)) * V; b_k[r] = t / (TILE / V) + r * (NT * V / TILE); }
    else      { b_k[r] = (t % (BK / V)) * V;   b_n[r] = t / (BK / V) + r * (NT * V / BK); }
    const int m = tile_m + a_m[r], n = tile_n + b_n[r];
    a_ok[r] = m < M;
    b_ok[r] = n < N;
    a_off[r] = a_ok[r] ? off2(m, d.M0, d.am0, d.am1) : 0;
    b_off[r] = b_ok[r] ? off2(n, d.N0, d.bn0, d.bn1) : 0;
    a_kp[r].init(kbeg + a_k[r], d.K0, d.ak0, d.ak1);
    b_kp[r].init(kbeg + b_k[r], d.K0, d.bk0, d.bk1);
  }
  // loads of one K-tile (live: a K-tile inside [kbeg, K)).  A slot outside the output tile, or any slot of a
  // dead tile, reads the operand's first vector instead, so the loop body stays free of branches: such
  // values reach only accumulator rows (columns) past M (N), which are never stored, or the LDS buffer
  // the last iteration fills and nobody reads.
  auto gload = [&](VT (&xa)[NVEC], VT (&xb)[NVEC], bool live) {
#pragma unroll
    for (int r = 0; r < NVEC; ++r) {
      const bool oka = live && a_ok[r], okb = live && b_ok[r];
      xa[r] = *reinterpret_cast<const VT*>(A + (oka ? a_off[r] + a_kp[r].off : 0));
      xb[r] = *reinterpret_cast<const VT*>(B + (okb ? b_off[r] + b_kp[r].off : 0));
      // one K0 run holds whole K-tiles: at most one carry per tile
      a_kp[r].k0 += BK;
      a_kp[r].off += (long long)BK * d.ak0;
      if (a_kp[r].k0 >= d.K0) { a_kp[r].k0 -= d.K0; a_kp[r].off += a_dk; }
      b_kp[r].k0 += BK;
      b_kp[r].off += (long long)BK * d.bk0;
      if (b_kp[r].k0 >= d.K0) { b_kp[r].k0 -= d.K0; b_kp[r].off += b_dk; }
    }
  };
  auto stage = [&](const VT (&xa)[NVEC], const VT (&xb)[NVEC], int buf) {
#pragma unroll
    for (int r = 0; r < NVEC; ++r) {
      if (A_MV) *reinterpret_cast<VT*>(&As[buf][a_k[r]][a_m[r]]) = xa[r];
      else {
#pragma unroll
        for (int e = 0; e < V; ++e) As[buf][a_k[r] + e][a_m[r]] = vec_elem(xa[r], e);
      }
      if (B_NV) *reinterpret_cast<VT*>(&Bs[buf][b_k[r]][b_n[r]]) = xb[r];
      else {
#pragma unroll
        for (int e = 0; e < V; ++e) Bs[buf][b_k[r] + e][b_n[r]] = vec_elem(xb[r], e);
      }
    }
  };

  typedef typename MfmaOp<T>::acc_t acc_t;
  acc_t acc[FB][FB];
#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  // One K-tile per iteration: its fragments from LDS buffer `buf`, the first half of its MFMAs, the next
  // K-tile (in registers since the previous iteration) stored into the other buffer, the second half
  // of the MFMAs, the loads of the K-tile after it, one barrier.  The store of the last iteration goes
  // to the buffer nobody reads again, and its loads are dead (they read the operands' first vector).
  VT ra[NVEC], rb[NVEC];
  gload(ra, rb, true);
  stage(ra, rb, 0);
  gload(ra, rb, kbeg + BK < K);
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < K; k0 += BK, buf ^= 1) {
    // one k-step (4 of K) of MFMAs from the fragments of LDS buffer `buf`
    auto kstep = [&](int kk) {
      T af[FB], bf[FB];
#pragma unroll
      for (int i = 0; i < FB; ++i) af[i] = As[buf][4 * kk + (lane >> 4)][wm * (TILE / 2) + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FB; ++j) bf[j] = Bs[buf][4 * kk + (lane >> 4)][wn * (TILE / 2) + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FB; ++i)
#pragma unroll
        for (int j = 0; j < FB; ++j) acc[i][j] = MfmaOp<T>::run(af[i], bf[j], acc[i][j]);
    };
#pragma unroll
    for (int kk = 0; kk < BK / 8; ++kk) kstep(kk);
    stage(ra, rb, buf ^ 1);
    // the next loads reuse the staged registers: keep them after the stores (hoisted above, they would
    // need a second register set and a copy that waits for them within the same K-tile)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int kk = BK / 8; kk < BK / 4; ++kk) kstep(kk);
    gload(ra, rb, k0 + 2 * BK < K);
    __syncthreads();
  }

  const T alpha = (T)d.alpha, beta = (T)d.beta;
  T* Wz = Wsp ? Wsp + (size_t)blockIdx.z * M * N : nullptr;
#pragma unroll
  for (int j = 0; j < FB; ++j) {
    int n = tile_n + wn * (TILE / 2) + j * 16 + (lane & 15);
    if (n >= N) continue;
    long long noff = off2(n, d.N0, d.cn0, d.cn1);
#pragma unroll
    for (int i = 0; i < FB; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = tile_m + wm * (TILE / 2) + i * 16 + MfmaOp<T>::row(lane, r);
        if (m >= M) continue;
        if (Wz) {
          Wz[m + (size_t)M * n] = (T)acc[i][j][r];
          continue;
        }
        long long o = noff + off2(m, d.M0, d.cm0, d.cm1);
        T v = alpha * (T)acc[i][j][r];
        if (d.beta != 0.0) v += beta * C[o];
        C[o] = v;
      }
    }
  }
}

// split-K reduction: C = alpha * sum_s W[s] + beta * C over the strided output (slices summed in
// order, so the result does not depend on scheduling)
// The grid's y index walks the output columns n and x the rows m, so no element index is divided (the
// first version split a 64-bit element index by M per element, and its integer divisions, not the
// memory traffic, set its time); the column offset is formed once per column, the row offset without a
// division when the row index is not split (M1 = 1).
template <typename T>
__global__ __launch_bounds__(NT) void tc_splitk_reduce(TcDesc d, const T* __restrict__ Wsp, int nsplit,
                                                       T* __restrict__ C) {
  const int M = d.M0 * d.M1, N = d.N0 * d.N1;
  const size_t MN = (size_t)M * N;
  for (int n = blockIdx.y; n < N; n += gridDim.y) {
    const long long on = off2(n, d.N0, d.cn0, d.cn1);
    for (int m = blockIdx.x * NT + threadIdx.x; m < M; m += gridDim.x * NT) {
      const size_t e = (size_t)m + (size_t)M * n;
      double acc = 0.0;
      for (int s = 0; s < nsplit; ++s) acc += (double)Wsp[e + (size_t)s * MN];
      const long long o = (d.M1 == 1 ? (long long)m * d.cm0 : off2(m, d.M0, d.cm0, d.cm1)) + on;
      double v = d.alpha * acc;
      if (d.beta != 0.0) v += d.beta * (double)C[o];
      C[o] = (T)v;
    }
  }
}

// ---------------------------------------------------------------------------
// BF16 kernel: fp32 operands in HBM rounded to bf16 while staging, m-major LDS
// tiles so each lane fetches its 4 consecutive-k bf16 values in one 8-byte read.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned int u = __float_as_uint(x);
  // round to nearest even (NaN stays NaN: quiet bit kept by the shift)
  unsigned int lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (unsigned short)(u >> 16);
}

__global__ __launch_bounds__(NT) void tc_kernel_bf16(TcDesc d, const float* __restrict__ A,
                                                     const float* __restrict__ B,
                                                     float* __restrict__ C, float* __restrict__ Wsp) {
  constexpr int BK = 32;
  constexpr int KP = BK + 4;   // row pitch in bf16 elements (8-byte aligned rows)
  __shared__ unsigned short As[BM][KP];
  __shared__ unsigned short Bs[BN][KP];

  const int M = d.M0 * d.M1, N = d.N0 * d.N1;
  const int tile_m = blockIdx.x * BM, tile_n = blockIdx.y * BN;
  const int bz = Wsp ? 0 : blockIdx.z;
  const int kbeg = Wsp ? blockIdx.z * d.kchunk : 0;
  const int K = Wsp ? min(d.K0 * d.K1, kbeg + d.kchunk) : d.K0 * d.K1;
  A += (long long)bz * d.abat;
  B += (long long)bz * d.bbat;
  C += (long long)bz * d.cbat;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // each thread stages 8 elements of A and 8 of B per K-tile (64x32 tiles)
  const bool a_mfast = (d.am0 == 1 && d.M0 > 1) || d.ak0 != 1;
  int a_mi[8], a_ki[8];
  long long a_moff[8];
  bool a_mok[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    if (a_mfast) { a_mi[r] = t & 63; a_ki[r] = (t >> 6) + 4 * r; }
    else         { a_ki[r] = t & 31; a_mi[r] = (t >> 5) + 8 * r; }
    int m = tile_m + a_mi[r];
    a_mok[r] = m < M;
    a_moff[r] = a_mok[r] ? off2(m, d.M0, d.am0, d.am1) : 0;
  }
  const bool b_nfast = (d.bn0 == 1 && d.N0 > 1) || d.bk0 != 1;
  int b_ni[8], b_ki[8];
  long long b_noff[8];
  bool b_nok[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    if (b_nfast) { b_ni[r] = t & 63; b_ki[r] = (t >> 6) + 4 * r; }
    else         { b_ki[r] = t & 31; b_ni[r] = (t >> 5) + 8 * r; }
    int n = tile_n + b_ni[r];
    b_nok[r] = n < N;
    b_noff[r] = b_nok[r] ? off2(n, d.N0, d.bn0, d.bn1) : 0;
  }
  float ra[8], rb[8];
  KPos a_kp[8], b_kp[8];
  const long long a_dk = d.ak1 - (long long)d.K0 * d.ak0, b_dk = d.bk1 - (long long)d.K0 * d.bk0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    a_kp[r].init(kbeg + a_ki[r], d.K0, d.ak0, d.ak1);
    b_kp[r].init(kbeg + b_ki[r], d.K0, d.bk0, d.bk1);
  }
  // loads the K-tile starting at k0, then moves the loaders one tile on
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int k = k0 + a_ki[r];
      float v = 0.f;
      if (a_mok[r] && k < K) v = A[a_moff[r] + a_kp[r].off];
      ra[r] = v;
      int kb = k0 + b_ki[r];
      float w = 0.f;
      if (b_nok[r] && kb < K) w = B[b_noff[r] + b_kp[r].off];
      rb[r] = w;
      a_kp[r].advance(BK, d.K0, d.ak0, a_dk);
      b_kp[r].advance(BK, d.K0, d.bk0, b_dk);
    }
  };

  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0, 0, 0, 0};

  gload(kbeg);
  for (int k0 = kbeg; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      As[a_mi[r]][a_ki[r]] = f2bf(ra[r]);
      Bs[b_ni[r]][b_ki[r]] = f2bf(rb[r]);
    }
    __syncthreads();
    if (k0 + BK < K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      s4 af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        af[i] = *reinterpret_cast<const s4*>(&As[wm * 32 + i * 16 + (lane & 15)][kk + 4 * (lane >> 4)]);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        bf[j] = *reinterpret_cast<const s4*>(&Bs[wn * 32 + j * 16 + (lane & 15)][kk + 4 * (lane >> 4)]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
  }
  const float alpha = (float)d.alpha, beta = (float)d.beta;
  float* Wz = Wsp ? Wsp + (size_t)blockIdx.z * M * N : nullptr;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    int n = tile_n + wn * 32 + j * 16 + (lane & 15);
    if (n >= N) continue;
    long long noff = off2(n, d.N0, d.cn0, d.cn1);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = tile_m + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        if (m >= M) continue;
        if (Wz) {
          Wz[m + (size_t)M * n] = acc[i][j][r];
          continue;
        }
        long long o = noff + off2(m, d.M0, d.cm0, d.cm1);
        float v = alpha * acc[i][j][r];
        if (d.beta != 0.0) v += beta * C[o];
        C[o] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// BF16 kernel on gfx950's double-K matrix instruction (v_mfma_f32_16x16x32_bf16: eight consecutive k per
// lane and operand, FP32 accumulation), for the layouts the 16-byte-load kernel accepts: FP32 operands
// read as float4 vectors (launch_vec's rules, K-tile 32), rounded to BF16 (nearest even) while they are
// stored into m-major (n-major) LDS tiles, so that each lane fetches its eight k values of a fragment with
// one 16-byte read.  64 x 64 output tile, 4 waves in 2 x 2, double-buffered LDS, one barrier per K-tile,
// branch-free loads as in tc_kernel_vec.
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <bool A_MV, bool B_NV>
__global__ __launch_bounds__(NT, 4) void tc_kernel_vec_bf16(TcDesc d, const float* __restrict__ A,
                                                            const float* __restrict__ B, float* __restrict__ C,
                                                            float* __restrict__ Wsp) {
  constexpr int TILE = 64, BK = 32, V = 4, KP = BK + 8;   // KP: row pitch in bf16 (80 bytes)
  constexpr int NVEC = TILE * BK / (NT * V);               // float4 vectors per thread and operand
  constexpr int FB = TILE / 32;
  static_assert(NVEC == 2, "loader coverage");
  __shared__ __attribute__((aligned(16))) unsigned short As[2][TILE][KP];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][TILE][KP];

  const int M = d.M0 * d.M1, N = d.N0 * d.N1;
  const int tile_m = blockIdx.x * TILE, tile_n = blockIdx.y * TILE;
  const int bz = Wsp ? 0 : blockIdx.z;
  const int kbeg = Wsp ? blockIdx.z * d.kchunk : 0;
  const int K = Wsp ? min(d.K0 * d.K1, kbeg + d.kchunk) : d.K0 * d.K1;
  A += (long long)bz * d.abat;
  B += (long long)bz * d.bbat;
  C += (long long)bz * d.cbat;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  int a_m[NVEC], a_k[NVEC], b_n[NVEC], b_k[NVEC];
  long long a_off[NVEC], b_off[NVEC];
  bool a_ok[NVEC], b_ok[NVEC];
  KPos a_kp[NVEC], b_kp[NVEC];
  const long long a_dk = d.ak1 - (long long)d.K0 * d.ak0, b_dk = d.bk1 - (long long)d.K0 * d.bk0;
#pragma unroll
  for (int r = 0; r < NVEC; ++r) {
    if (A_MV) { a_m[r] = (t % (TILE / V)) * V; a_k[r] = t / (TILE / V) + r * (NT * V / TILE); }
    else      { a_k[r] = (t % (BK / V)) * V;   a_m[r] = t / (BK / V) + r * (NT * V / BK); }
    if (B_NV) { b_n[r] = (t % (TILE / V)) * V; b_k[r] = t / (TILE / V) + r * (NT * V / TILE); }
    else      { b_k[r] = (t % (BK / V)) * V;   b_n[r] = t / (BK / V) + r * (NT * V / BK); }
    const int m = tile_m + a_m[r], n = tile_n + b_n[r];
    a_ok[r] = m < M;
    b_ok[r] = n < N;
    a_off[r] = a_ok[r] ? off2(m, d.M0, d.am0, d.am1) : 0;
    b_off[r] = b_ok[r] ? off2(n, d.N0, d.bn0, d.bn1) : 0;
    a_kp[r].init(kbeg + a_k[r], d.K0, d.ak0, d.ak1);
    b_kp[r].init(kbeg + b_k[r], d.K0, d.bk0, d.bk1);
  }
  auto gload = [&](f4 (&xa)[NVEC], f4 (&xb)[NVEC], bool live) {
#pragma unroll
    for (int r = 0; r < NVEC; ++r) {
      const bool oka = live && a_ok[r], okb = live && b_ok[r];
      xa[r] = *reinterpret_cast<const f4*>(A + (oka ? a_off[r] + a_kp[r].off : 0));
      xb[r] = *reinterpret_cast<const f4*>(B + (okb ? b_off[r] + b_kp[r].off : 0));
      a_kp[r].k0 += BK;
      a_kp[r].off += (long long)BK * d.ak0;
      if (a_kp[r].k0 >= d.K0) { a_kp[r].k0 -= d.K0; a_kp[r].off += a_dk; }
      b_kp[r].k0 += BK;
      b_kp[r].off += (long long)BK * d.bk0;
      if (b_kp[r].k0 >= d.K0) { b_kp[r].k0 -= d.K0; b_kp[r].off += b_dk; }
    }
  };
  // a vector along k becomes one 8-byte store of four BF16 values; a vector along m (n) four 2-byte ones
  auto put = [&](unsigned short (&T)[TILE][KP], int row, int k, const f4& v, bool along_rows) {
    if (along_rows) {
#pragma unroll
      for (int e = 0; e < V; ++e) T[row + e][k] = f2bf(v[e]);
    } else {
      const unsigned int lo = (unsigned int)f2bf(v[0]) | ((unsigned int)f2bf(v[1]) << 16);
      const unsigned int hi = (unsigned int)f2bf(v[2]) | ((unsigned int)f2bf(v[3]) << 16);
      *reinterpret_cast<uint2*>(&T[row][k]) = make_uint2(lo, hi);
    }
  };
  auto stage = [&](const f4 (&xa)[NVEC], const f4 (&xb)[NVEC], int buf) {
#pragma unroll
    for (int r = 0; r < NVEC; ++r) {
      put(As[buf], a_m[r], a_k[r], xa[r], A_MV);
      put(Bs[buf], b_n[r], b_k[r], xb[r], B_NV);
    }
  };

  f4 acc[FB][FB];
#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) acc[i][j] = f4{0, 0, 0, 0};
  // one k-step (32 of K): lane l supplies rows (columns) l % 16 of each 16-block, k = 8 (l / 16) .. + 7
  auto kstep = [&](int buf) {
    bf8 af[FB], bf[FB];
#pragma unroll
    for (int i = 0; i < FB; ++i)
      af[i] = *reinterpret_cast<const bf8*>(&As[buf][wm * (TILE / 2) + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
    for (int j = 0; j < FB; ++j)
      bf[j] = *reinterpret_cast<const bf8*>(&Bs[buf][wn * (TILE / 2) + j * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
    for (int i = 0; i < FB; ++i)
#pragma unroll
      for (int j = 0; j < FB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
  };

  f4 ra[NVEC], rb[NVEC];
  gload(ra, rb, true);
  stage(ra, rb, 0);
  gload(ra, rb, kbeg + BK < K);
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < K; k0 += BK, buf ^= 1) {
    kstep(buf);
    stage(ra, rb, buf ^ 1);
    __builtin_amdgcn_sched_barrier(0);
    gload(ra, rb, k0 + 2 * BK < K);
    __syncthreads();
  }

  const float alpha = (float)d.alpha, beta = (float)d.beta;
  float* Wz = Wsp ? Wsp + (size_t)blockIdx.z * M * N : nullptr;
#pragma unroll
  for (int j = 0; j < FB; ++j) {
    int n = tile_n + wn * (TILE / 2) + j * 16 + (lane & 15);
    if (n >= N) continue;
    long long noff = off2(n, d.N0, d.cn0, d.cn1);
#pragma unroll
    for (int i = 0; i < FB; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = tile_m + wm * (TILE / 2) + i * 16 + 4 * (lane >> 4) + r;
        if (m >= M) continue;
        if (Wz) {
          Wz[m + (size_t)M * n] = acc[i][j][r];
          continue;
        }
        long long o = noff + off2(m, d.M0, d.cm0, d.cm1);
        float v = alpha * acc[i][j][r];
        if (d.beta != 0.0) v += beta * C[o];
        C[o] = v;
      }
    }
  }
}

}  // namespace

static size_t extent_bytes(long long x0, long long s0, long long x1, long long s1, long long y0, long long t0,
                           long long y1, long long t1, long long nb, long long sb, size_t es) {
  if (s0 < 0 || s1 < 0 || t0 < 0 || t1 < 0 || sb < 0) throw Error(TNAD_ERR_ARG, "tc: negative stride");
  long long last = (x0 - 1) * s0 + (x1 - 1) * s1 + (y0 - 1) * t0 + (y1 - 1) * t1 + (nb - 1) * sb;
  return (size_t)(last + 1) * es;
}

static bool tc_vec_disabled() {
  static const bool off = [] {
    const char* e = std::getenv("TNAD_TC_NOVEC");
    return e && *e && *e != '0';
  }();
  return off;
}

// Vector layout of one operand for tc_kernel_vec (V elements of 16 bytes): 1 = along its M (N) index,
// 2 = along K, 0 = none.  X0, X1 / x0, x1: extents and strides of the operand's M (N) sub-indices; K0, K1 /
// k0, k1 those of K.  Every vector must be V consecutive elements of one sub-index run starting at a
// multiple of V elements from a 16-byte aligned base.
static int vec_mode(int X0, int X1, long long x0, long long x1, int K0, int K1, long long k0, long long k1,
                    int batch, long long bat, const void* p, int V) {
  auto mult = [&](long long stride, int ext) { return ext <= 1 || stride % V == 0; };
  if (reinterpret_cast<uintptr_t>(p) % 16 != 0 || !mult(bat, batch)) return 0;
  if (X0 > 1 && x0 == 1 && X0 % V == 0 && mult(x1, X1) && mult(k0, K0) && mult(k1, K1)) return 1;
  if (K0 > 1 && k0 == 1 && K0 % V == 0 && mult(k1, K1) && mult(x0, X0) && mult(x1, X1)) return 2;
  return 0;
}

// BF16 operands on the double-K instruction where launch_vec's FP32 rules (K-tile 32) hold
static bool launch_vec_bf16(const TcDesc& dd, dim3 grid64, const void* A, const void* B, void* C, void* W,
                            hipStream_t s) {
  constexpr int V = 4;
  const long long K = (long long)dd.K0 * dd.K1;
  if (K % 32 != 0 || (dd.K1 > 1 && dd.K0 % 32 != 0)) return false;
  const int ma = vec_mode(dd.M0, dd.M1, dd.am0, dd.am1, dd.K0, dd.K1, dd.ak0, dd.ak1, dd.batch, dd.abat, A, V);
  const int mb = vec_mode(dd.N0, dd.N1, dd.bn0, dd.bn1, dd.K0, dd.K1, dd.bk0, dd.bk1, dd.batch, dd.bbat, B, V);
  if (ma == 0 || mb == 0) return false;
  const float* a = (const float*)A;
  const float* b = (const float*)B;
  float* c = (float*)C;
  float* w = (float*)W;
  if (ma == 1 && mb == 1) hipLaunchKernelGGL((tc_kernel_vec_bf16<true, true>), grid64, dim3(NT), 0, s, dd, a, b, c, w);
  else if (ma == 1) hipLaunchKernelGGL((tc_kernel_vec_bf16<true, false>), grid64, dim3(NT), 0, s, dd, a, b, c, w);
  else if (mb == 1) hipLaunchKernelGGL((tc_kernel_vec_bf16<false, true>), grid64, dim3(NT), 0, s, dd, a, b, c, w);
  else hipLaunchKernelGGL((tc_kernel_vec_bf16<false, false>), grid64, dim3(NT), 0, s, dd, a, b, c, w);
  return true;
}

template <typename T, int TILE, int BK>
static bool launch_vec(const TcDesc& dd, dim3 grid, const void* A, const void* B, void* C, void* W, hipStream_t s) {
  constexpr int V = 16 / sizeof(T);
  const long long K = (long long)dd.K0 * dd.K1;
  // whole K-tiles, each inside one K0 run (split-K slices are multiples of 32)
  if (K % BK != 0 || (dd.K1 > 1 && dd.K0 % BK != 0)) return false;
  const int ma = vec_mode(dd.M0, dd.M1, dd.am0, dd.am1, dd.K0, dd.K1, dd.ak0, dd.ak1, dd.batch, dd.abat, A, V);
  const int mb = vec_mode(dd.N0, dd.N1, dd.bn0, dd.bn1, dd.K0, dd.K1, dd.bk0, dd.bk1, dd.batch, dd.bbat, B, V);
  if (ma == 0 || mb == 0) return false;
  const T* a = (const T*)A;
  const T* b = (const T*)B;
  T* c = (T*)C;
  T* w = (T*)W;
  if (ma == 1 && mb == 1) hipLaunchKernelGGL((tc_kernel_vec<T, TILE, BK, true, true>), grid, dim3(NT), 0, s, dd, a, b, c, w);
  else if (ma == 1) hipLaunchKernelGGL((tc_kernel_vec<T, TILE, BK, true, false>), grid, dim3(NT), 0, s, dd, a, b, c, w);
  else if (mb == 1) hipLaunchKernelGGL((tc_kernel_vec<T, TILE, BK, false, true>), grid, dim3(NT), 0, s, dd, a, b, c, w);
  else hipLaunchKernelGGL((tc_kernel_vec<T, TILE, BK, false, false>), grid, dim3(NT), 0, s, dd, a, b, c, w);
  return true;
}

void tc_launch(Ctx* c, Prec p, const TcDesc& d, const void* A, const void* B, void* C, hipStream_t s) {
  const long long M = (long long)d.M0 * d.M1, N = (long long)d.N0 * d.N1;
  {
    const size_t es = p == Prec::F64 ? 8 : 4;
    c->check_range(A, extent_bytes(d.M0, d.am0, d.M1, d.am1, d.K0, d.ak0, d.K1, d.ak1, d.batch, d.abat, es), "tc A");
    c->check_range(B, extent_bytes(d.K0, d.bk0, d.K1, d.bk1, d.N0, d.bn0, d.N1, d.bn1, d.batch, d.bbat, es), "tc B");
    c->check_range(C, extent_bytes(d.M0, d.cm0, d.M1, d.cm1, d.N0, d.cn0, d.N1, d.cn1, d.batch, d.cbat, es), "tc C");
  }
  if (M <= 0 || N <= 0 || d.batch <= 0) return;
  if ((long long)d.K0 * d.K1 <= 0) {
    throw Error(TNAD_ERR_ARG, "tc_launch: empty contraction index");
  }
  if (d.M0 <= 0 || d.M1 <= 0 || d.N0 <= 0 || d.N1 <= 0 || d.K0 <= 0 || d.K1 <= 0 ||
      M > (1LL << 30) || N > (1LL << 30) || d.batch > 65535)
    throw Error(TNAD_ERR_ARG, "tc_launch: extents out of range");
  dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + BN - 1) / BN), (unsigned)d.batch);
  if (grid.y > 65535) throw Error(TNAD_ERR_ARG, "tc_launch: N too large");
  const long long K = (long long)d.K0 * d.K1;
  const long long tiles64 = (long long)grid.x * grid.y;
  // 32 x 32 tiles (4x the workgroups) when 64 x 64 tiles would leave most CUs idle: short K, or an
  // output of at most 16 tiles with K up to 4096 (the k x k Gram products of the eigensolver, cut along
  // K below as well; for longer K the 64 x 64 tile's higher MFMA rate per slice wins)
  const dim3 grid32((unsigned)((M + 31) / 32), (unsigned)((N + 31) / 32), (unsigned)d.batch);
  const bool small = (long long)grid.x * grid.y * grid.z < TC_SMALL_BELOW && grid32.y <= 65535 &&
                     (K < 256 || (tiles64 <= 16 && K <= 4096));
  if (small) grid = grid32;
  // split-K when the output tiles cannot fill the chip and K is long: about 512 workgroups,
  // slices of at least 128 K elements (a multiple of both kernels' K-tile)
  const long long tiles = (long long)grid.x * grid.y;
  TcDesc dd = d;
  int nsplit = 1;
  if (d.batch == 1 && tiles < (small ? 512 : 128) && K >= 256) {
    long long want = std::min<long long>((512 + tiles - 1) / tiles, (K + 127) / 128);
    long long chunk = (K + want - 1) / want;
    chunk = (chunk + 31) / 32 * 32;
    nsplit = (int)((K + chunk - 1) / chunk);
    dd.kchunk = (int)chunk;
  }
  DevBuf wsp;
  if (nsplit > 1) {
    const size_t es = p == Prec::F64 ? 8 : 4;
    wsp = DevBuf(c, (size_t)nsplit * M * N * es);
    grid.z = (unsigned)nsplit;
  }
  void* W = nsplit > 1 ? wsp.ptr : nullptr;
  // 128 x 128 tiles for large outputs (at least TC_BIG_MIN_WG of them); 64 x 64 otherwise
  const dim3 grid_big((unsigned)((M + 127) / 128), (unsigned)((N + 127) / 128), grid.z);
  const bool big = !small && nsplit == 1 && (long long)grid_big.x * grid_big.y * grid_big.z >= TC_BIG_MIN_WG;
  const bool deep = !big && !small && (long long)grid.x * grid.y * grid.z < TC_DEEP_BELOW;
  // The 16-byte-load kernel where the operand layouts allow it: 128 x 128 tiles for large outputs, as
  // above (a wave then owns a 64 x 64 block: four times the MFMA work per fragment read and per barrier
  // of the 64 x 64 tile; 4096^3 FP64 35.5 -> 62.6 TFLOP/s, FP32 84.5 -> 124.3; at one workgroup per CU
  // the 64 x 64 tile is still faster), 32 x 32 for small grids, 64 x 64 otherwise
  bool vec = false;
  const bool vbig = big;
  if (!tc_vec_disabled()) {
    if (p == Prec::F64)
      vec = vbig ? launch_vec<double, 128, 16>(dd, grid_big, A, B, C, W, s)
                 : small ? launch_vec<double, 32, 32>(dd, grid, A, B, C, W, s)
                         : !big && launch_vec<double, 64, 16>(dd, grid, A, B, C, W, s);
    else if (p == Prec::F32)
      vec = vbig ? launch_vec<float, 128, 16>(dd, grid_big, A, B, C, W, s)
                 : small ? launch_vec<float, 32, 32>(dd, grid, A, B, C, W, s)
                         : !big && launch_vec<float, 64, 16>(dd, grid, A, B, C, W, s);
    else
      vec = launch_vec_bf16(dd, dim3((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), grid.z), A, B, C, W, s);
  }
  if (!vec) switch (p) {
    case Prec::F64:
      if (small)
        hipLaunchKernelGGL((tc_kernel<double, TC_BK, 32>), grid, dim3(NT), 0, s, dd, (const double*)A,
                           (const double*)B, (double*)C, (double*)W);
      else if (big)
        hipLaunchKernelGGL((tc_kernel<double, TC_BK, 128>), grid_big, dim3(NT), 0, s, dd, (const double*)A,
                           (const double*)B, (double*)C, (double*)W);
      else if (deep)
        hipLaunchKernelGGL((tc_kernel<double, TC_BK_DEEP, 64>), grid, dim3(NT), 0, s, dd, (const double*)A,
                           (const double*)B, (double*)C, (double*)W);
      else
        hipLaunchKernelGGL((tc_kernel<double, TC_BK, 64, 2>), grid, dim3(NT), 0, s, dd, (const double*)A,
                           (const double*)B, (double*)C, (double*)W);
      break;
    case Prec::F32:
      if (small)
        hipLaunchKernelGGL((tc_kernel<float, TC_BK, 32>), grid, dim3(NT), 0, s, dd, (const float*)A,
                           (const float*)B, (float*)C, (float*)W);
      else if (big)
        hipLaunchKernelGGL((tc_kernel<float, TC_BK, 128>), grid_big, dim3(NT), 0, s, dd, (const float*)A,
                           (const float*)B, (float*)C, (float*)W);
      else if (deep)
        hipLaunchKernelGGL((tc_kernel<float, TC_BK_DEEP, 64>), grid, dim3(NT), 0, s, dd, (const float*)A,
                           (const float*)B, (float*)C, (float*)W);
      else
        hipLaunchKernelGGL((tc_kernel<float, TC_BK, 64, 2>), grid, dim3(NT), 0, s, dd, (const float*)A,
                           (const float*)B, (float*)C, (float*)W);
      break;
    case Prec::BF16:
      hipLaunchKernelGGL(tc_kernel_bf16, grid, dim3(NT), 0, s, dd, (const float*)A,
                         (const float*)B, (float*)C, (float*)W);
      break;
  }
  TNAD_HIP_TRY(hipGetLastError());
  if (nsplit > 1) {
    if (M * N >= (1LL << 31)) throw Error(TNAD_ERR_ARG, "split-K output too large");
    const dim3 rg((unsigned)std::min<long long>((M + NT - 1) / NT, 65535LL),
                  (unsigned)std::min<long long>(N, 65535LL));
    if (p == Prec::F64)
      hipLaunchKernelGGL(tc_splitk_reduce<double>, rg, dim3(NT), 0, s, dd, (const double*)W, nsplit,
                         (double*)C);
    else
      hipLaunchKernelGGL(tc_splitk_reduce<float>, rg, dim3(NT), 0, s, dd, (const float*)W, nsplit,
                         (float*)C);
    TNAD_HIP_TRY(hipGetLastError());
  }
}

}  // namespace tnad
