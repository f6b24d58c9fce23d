// Small dense FP64 kernels for the eigensolver, one workgroup each (m <= 192).
//
//  * jacobi_rot_kernel / jacobi_apply_kernel : cyclic two-sided Jacobi with the
//    round-robin (circle) ordering on the upper triangle of 2x2 blocks in LDS
//    (192x192 FP64 = 149 KB of the 160 KB a CU offers), so each round applies
//    m/2 disjoint rotations in parallel and every 2x2 block is updated by one
//    thread; the rotations are recorded and replayed on the identity by many
//    workgroups.
//    Used for the full spectrum when n <= 192 and for the Rayleigh-Ritz problem
//    of the block subspace iteration.
//  * ritz_select_kernel : Rayleigh-Ritz of a nearly diagonal matrix: a few Jacobi
//    rotations for the strongly coupled pairs, perturbation theory for the rest.
//  * chol_inv_kernel : Cholesky of a (shifted) Gram matrix followed by the
//    triangular inverse, register-resident, for CholeskyQR orthonormalisation.
#include "tnad_internal.h"

namespace tnad {
namespace {

constexpr int EIG_NT = 1024;
// Jacobi rotation kernel: JR_NT threads, each holding up to JB_MAX 2x2 blocks (and their fixed
// addresses) in registers: ceil(P (P + 1) / 2 / JR_NT) for P = JACOBI_MAX / 2.
constexpr int JR_NT = 1024;
constexpr int JB_MAX = ((JACOBI_MAX / 2) * (JACOBI_MAX / 2 + 1) / 2 + JR_NT - 1) / JR_NT;

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

__device__ __forceinline__ void rr_pair(int r_in, int t, int M, int& p, int& q) {
  // round-robin (circle) pairing of round r_in (0 <= r_in < M = mm - 1), slot t (0 <= t < mm / 2):
  // slot 0 pairs r_in with the fixed index M, slot t > 0 pairs r_in + t with r_in - t (mod M)
  if (t == 0) { p = r_in; q = M; }
  else {
    p = r_in + t;
    if (p >= M) p -= M;
    q = r_in - t;
    if (q < 0) q += M;
  }
  if (p > q) { int tmp = p; p = q; q = tmp; }
}

// Jacobi sweeps on the matrix in LDS, stored in slot-position order.  Round r of the round-robin
// ordering pairs, in slot t, the indices a = r + t and b = r - t (mod M = mm - 1; slot 0 pairs r with
// the fixed index M).  Position 2t holds a and 2t + 1 holds b, so every round works on the same
// positions: a thread's 2x2 blocks (pairs of slots) sit at addresses fixed for the whole run.
// Between rounds the indices move by a fixed permutation of positions (a-members one slot down,
// b-members one slot up, around the ends; see jpos_next), which the rotated blocks follow on
// write-back; after a full sweep (M rounds) every index is back where it started.  An odd m is
// padded to even mm with a zero row and column (never rotated).
//
// Layout: the upper triangle of 2x2 blocks (slot i <= slot j, block number jblk(i, j) =
// j (j + 1) / 2 + i) in four planes of P (P + 1) / 2 doubles, one per block element (a_i a_j, a_i b_j,
// b_i a_j, b_i b_j).  Threads own consecutive blocks, so each plane is read with consecutive 8-byte
// words across a wavefront, and most write-backs (to blocks (i -+ 1, j -+ 1)) stay consecutive too.
// A diagonal block keeps its off-diagonal entry in the a_i b_j plane only; its b_i a_j word is
// never written and never read.
//
// The rotations of every round are recorded for the index pair (p, q) = (min, max) of each slot
// (rot[round * P + slot] = {c, s}); the eigenvector matrix is formed afterwards by
// jacobi_apply_kernel.  Writes the (unsorted) diagonal to diag and the number of rounds to
// *rounds_out.
__device__ __forceinline__ int jblk(int i, int j) { return j * (j + 1) / 2 + i; }
// plane address of the entry at positions (x, y): symmetric, so the pair is ordered by (slot, member)
__device__ __forceinline__ int jplane_adr(int x, int y, int nblk) {
  if (x > y) { const int tmp = x; x = y; y = tmp; }
  return (2 * (x & 1) + (y & 1)) * nblk + jblk(x >> 1, y >> 1);
}
__device__ __forceinline__ int jpos_next(int pos, int P) {   // position of an index one round later
  if (P == 1) return pos;                                    // one slot: the pair never changes
  const int t = pos >> 1;
  if ((pos & 1) == 0) return t == 0 ? 3 : (t == 1 ? 0 : pos - 2);   // a-members move down
  if (t == 0) return 1;                                              // the fixed index M
  return t == P - 1 ? pos - 1 : pos + 2;                             // b-members move up
}
__device__ __forceinline__ int jpos_index0(int pos, int M) {   // index at a position in round 0
  const int t = pos >> 1;
  if (pos == 0) return 0;
  if (pos == 1) return M;
  return (pos & 1) ? M - t : t;
}
__device__ __forceinline__ int jpos_prev(int pos, int P) {   // position of an index one round earlier
  if (P == 1) return pos;
  if ((pos & 1) == 0) return pos == 0 ? 2 : (pos == 2 * P - 2 ? 2 * P - 1 : pos + 2);
  return pos == 1 ? 1 : (pos == 3 ? 0 : pos - 2);
}

constexpr int JR_MAX_SWEEPS = 64;

// Rotation of one slot from its 2x2 diagonal block; returns false (c = 1, s = 0) below the thresholds.
// tan of the rotation angle (|angle| <= pi/4): tt = sign(d) e / (|d| + hypot(d, e)) with d = abb - aaa,
// e = 2 aab; then c = 1 / sqrt(1 + tt^2), s = tt c, formed without tt.
// whether the pair (a, b) with entries aaa, aab, abb rotates: |aab| above the matrix's rounding level and
// above 1e-15 of the geometric mean of the two diagonal entries
__device__ __forceinline__ bool jrot_needed(double aaa, double aab, double abb, double abs_floor) {
  const double rel_tol2 = 1e-30;   // (1e-15)^2
  return fabs(aab) > abs_floor && aab * aab > rel_tol2 * fabs(aaa * abb);
}
__device__ __forceinline__ bool jrot_cs(double aaa, double aab, double abb, double abs_floor, double& c,
                                        double& sn) {
  c = 1.0;
  sn = 0.0;
  if (!jrot_needed(aaa, aab, abb, abs_floor)) return false;
  const double d = abb - aaa, e = 2.0 * aab;
  const double den = fabs(d) + sqrt(d * d + e * e);
  const double inv = 1.0 / sqrt(den * den + e * e);
  c = den * inv;
  sn = (d >= 0.0 ? e : -e) * inv;
  return true;
}

// A' = Ji^T A Jj on a 2x2 block (a_i a_j, a_i b_j, b_i a_j, b_i b_j).  A diagonal block (i == j)
// holds its off-diagonal entry in x[1] only; its rotated off-diagonal entry is set to zero.
__device__ __forceinline__ void jblock_update(double (&x)[4], double ci, double si, double cj, double sj,
                                              bool dgn) {
  const double x2 = dgn ? x[1] : x[2];
  const double ypr = ci * x[0] - si * x2, yps = ci * x[1] - si * x[3];
  const double yqr = si * x[0] + ci * x2, yqs = si * x[1] + ci * x[3];
  x[0] = cj * ypr - sj * yps;
  x[1] = dgn ? 0.0 : sj * ypr + cj * yps;
  x[2] = dgn ? 0.0 : cj * yqr - sj * yqs;
  x[3] = sj * yqr + cj * yqs;
}

// record of slot t's rotation in global round R, for the sorted index pair: (c, s) on (a, b) is
// (c, -s) on (b, a)
__device__ __forceinline__ void jrot_record(double* rot, int R, int t, int M, int P, double c, double sn) {
  const int r = R % M;
  int ia, ib;
  if (t == 0) { ia = r; ib = M; }
  else {
    ia = r + t;
    if (ia >= M) ia -= M;
    ib = r - t;
    if (ib < 0) ib += M;
  }
  reinterpret_cast<double2*>(rot)[(size_t)R * P + t] = make_double2(c, ia < ib ? sn : -sn);
}

// NEXT: round R + 1's rotations are formed during round R (two barriers per round) by threads that own
// one block fewer than the busiest; otherwise (three barriers per round) the first P threads form each
// round's rotations before its block updates.  NEXT pays where the slowest threads own JB_MAX blocks
// and the rotation threads JB_MAX - 1 (P (P + 1) / 2 > (JB_MAX - 1) JR_NT), so the rotation work hides
// behind their block updates.
template <bool NEXT>
__global__ __launch_bounds__(JR_NT) void jacobi_rot_kernel(const double* __restrict__ H, int ldh, int m,
                                                            double* __restrict__ rot, int max_sweeps,
                                                            double* __restrict__ diag, int* rounds_out,
                                                            int* sweeps_out, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;   // the caller's result stands (see jacobi_eig_ws)
  extern __shared__ double smem[];
  const int mm = m + (m & 1);
  const int P = mm / 2;
  const int M = mm - 1;
  const int nblk = P * (P + 1) / 2;
  double* a = smem;                              // 4 planes of nblk blocks, position order
  __shared__ double2 rcs[2][JACOBI_MAX / 2];     // (cos, sin) per slot, this round and the next
  __shared__ double red[JR_NT / 64];
  __shared__ int nrot_sw[JR_MAX_SWEEPS];         // 1 if sweep k rotates at all
  __shared__ int4 pent_s[JACOBI_MAX / 2];        // next-round diagonal entries per rotation slot
  __shared__ double fro_s;
  const int t = threadIdx.x;
  for (int k = t; k < JR_MAX_SWEEPS; k += JR_NT) nrot_sw[k] = 0;

  // load (position order of round 0) and symmetrise (average both triangles so a slightly
  // asymmetric input is still handled)
  double part = 0.0;
  for (int q = t; q < 4 * nblk; q += JR_NT) {
    const int e = q / nblk, b = q - e * nblk;
    int i, j;
    decode_tri(b, i, j);
    const int x = jpos_index0(2 * i + (e >> 1), M), y = jpos_index0(2 * j + (e & 1), M);
    const double v = (x < m && y < m) ? 0.5 * (H[x + (long long)y * ldh] + H[y + (long long)x * ldh]) : 0.0;
    a[q] = v;
    // weight of the entry in |H|_F^2: an off-diagonal block holds four entries of the upper triangle;
    // a diagonal block two diagonal entries, one off-diagonal entry and its duplicate
    const double w = i != j ? 2.0 : (e == 1 ? 2.0 : (e == 2 ? 0.0 : 1.0));
    part += w * v * v;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o);
  if ((t & 63) == 0) red[t >> 6] = part;
  __syncthreads();
  if (t == 0) {
    double sacc = 0.0;
    for (int w = 0; w < JR_NT / 64; ++w) sacc += red[w];
    fro_s = sqrt(sacc);
  }
  __syncthreads();
  // An off-diagonal entry at or below the rounding level of the matrix (2 eps |H|_F) carries no
  // information: rotating it only reshuffles rounding noise, so sweeps would never end on matrices
  // whose spectrum spans many decades.  Above that, the relative criterion resolves small
  // eigenvalues to their own scale.
  const double abs_floor = 4.5e-16 * fro_s + 1e-300;
  // this thread's blocks b = t + u JR_NT: unordered slot pairs (i <= j), packed i | j << 16 (-1: none);
  // their entries are read from a[e nblk + b] and written one round later to the plane addresses of
  // the moved positions, packed two to an int (4 nblk < 2^16)
  int bij[JB_MAX], wadr[JB_MAX][2];
#pragma unroll
  for (int u = 0; u < JB_MAX; ++u) {
    const int b = t + u * JR_NT;
    int i = 0, j = 0;
    if (b < nblk) decode_tri(b, i, j);
    bij[u] = (b < nblk) ? (i | (j << 16)) : -1;
    int w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      w[e] = jplane_adr(jpos_next(2 * i + (e >> 1), P), jpos_next(2 * j + (e & 1), P), nblk);
    wadr[u][0] = w[0] | (w[1] << 16);
    wadr[u][1] = w[2] | (w[3] << 16);
  }
  // Rotations of round R + 1 are formed during round R by the last P threads (which own at most
  // JB_MAX - 1 blocks), from round R's blocks and rotations: slot s of round R + 1 pairs the indices
  // now at positions xo = jpos_prev(2 s) and yo = jpos_prev(2 s + 1), so its diagonal block is the
  // rotated entries (xo, xo), (xo, yo), (yo, yo) of this round.  Each entry is packed as block number
  // | element << 13 | slot of its row << 15 | slot of its column << 22 | diagonal block << 29 |
  // always zero << 30 (the off-diagonal entry of a diagonal block).
  const int ts = NEXT ? t - (JR_NT - P) : -1;   // rotation slot of this thread, or negative
  if (ts >= 0) {
    int pent[3];
    const int xo = jpos_prev(2 * ts, P), yo = jpos_prev(2 * ts + 1, P);
    const int pq[3][2] = {{xo, xo}, {xo, yo}, {yo, yo}};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      int x = pq[k][0], y = pq[k][1];
      if (x > y) { const int tmp = x; x = y; y = tmp; }
      const int i = x >> 1, j = y >> 1;
      const int zero = (i == j && x != y) ? 1 : 0;
      pent[k] = jblk(i, j) | ((2 * (x & 1) + (y & 1)) << 13) | (i << 15) | (j << 22) | ((i == j) << 29) |
                (zero << 30);
    }
    pent_s[ts] = make_int4(pent[0], pent[1], pent[2], 0);
  }

  // round 0's rotations
  if (NEXT && t < P) {
    const int bd = jblk(t, t);
    double c, sn;
    if (jrot_cs(a[bd], a[nblk + bd], a[3 * nblk + bd], abs_floor, c, sn)) nrot_sw[0] = 1;
    rcs[0][t] = make_double2(c, sn);
    jrot_record(rot, 0, t, M, P, c, sn);
  }
  __syncthreads();

  const int total_rounds = max_sweeps * M;
  int round = 0;
  int sweep = 0;
  int cur = 0;
  for (; sweep < max_sweeps; ++sweep) {
    for (int r = 0; r < M; ++r, ++round) {
      if (!NEXT) {
        if (t < P) {
          const int bd = jblk(t, t);
          double c, sn;
          if (jrot_cs(a[bd], a[nblk + bd], a[3 * nblk + bd], abs_floor, c, sn)) nrot_sw[sweep] = 1;
          rcs[0][t] = make_double2(c, sn);
          jrot_record(rot, round, t, M, P, c, sn);
        }
        __syncthreads();
      }
      if (ts >= 0 && round + 1 < total_rounds) {
        const int4 p4 = pent_s[ts];
        const int pent[3] = {p4.x, p4.y, p4.z};
        double v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int pe = pent[k];
          const int b = pe & 0x1fff, el = (pe >> 13) & 3, i = (pe >> 15) & 0x7f, j = (pe >> 22) & 0x7f;
          const bool dgn = (pe >> 29) & 1;
          double y[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) y[e] = a[e * nblk + b];
          const double2 csi = rcs[cur][i], csj = rcs[cur][j];
          jblock_update(y, csi.x, csi.y, csj.x, csj.y, dgn);
          v[k] = ((pe >> 30) & 1) ? 0.0 : (el == 0 ? y[0] : (el == 1 ? y[1] : (el == 2 ? y[2] : y[3])));
        }
        double c, sn;
        if (jrot_cs(v[0], v[1], v[2], abs_floor, c, sn)) nrot_sw[(round + 1) / M] = 1;
        rcs[cur ^ 1][ts] = make_double2(c, sn);
        jrot_record(rot, round + 1, ts, M, P, c, sn);
      }
      // 2x2 block updates A' = Ji^T A Jj over unordered pairs of rotation slots.  Every entry belongs
      // to exactly one block; all blocks are read, then (after a barrier, since the write-back goes
      // to next round's positions) written.
      double x[JB_MAX][4];
#pragma unroll
      for (int u = 0; u < JB_MAX; ++u) {
        const int b = bij[u] < 0 ? 0 : t + u * JR_NT;
#pragma unroll
        for (int e = 0; e < 4; ++e) x[u][e] = a[e * nblk + b];
      }
#pragma unroll
      for (int u = 0; u < JB_MAX; ++u) {
        const int bb = bij[u] < 0 ? 0 : bij[u];
        const double2 csi = rcs[cur][bb & 0xffff], csj = rcs[cur][bb >> 16];
        jblock_update(x[u], csi.x, csi.y, csj.x, csj.y, (bb & 0xffff) == (bb >> 16));
      }
      // the rotated entries are formed before the barrier: only they (not the entries and both
      // rotations of every block) stay live across it, which keeps the kernel within its registers
#pragma unroll
      for (int u = 0; u < JB_MAX; ++u) {
#pragma unroll
        for (int e = 0; e < 4; ++e) asm volatile("" : "+v"(x[u][e]));
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < JB_MAX; ++u) {
        if (bij[u] < 0) continue;
        // a diagonal block's two off-diagonal words are the same entry (zero): write it once
        const bool dgn = (bij[u] & 0xffff) == (bij[u] >> 16);
        a[wadr[u][0] & 0xffff] = x[u][0];
        a[(unsigned)wadr[u][0] >> 16] = x[u][1];
        if (!dgn) a[wadr[u][1] & 0xffff] = x[u][2];
        a[(unsigned)wadr[u][1] >> 16] = x[u][3];
      }
      __syncthreads();
      if (NEXT) cur ^= 1;
    }
    // every flag of this sweep was set before the last round's first barrier
    if (nrot_sw[sweep] == 0) { ++sweep; break; }
    // The next sweep would rotate no pair when every off-diagonal entry now fails the rotation test:
    // without rotations no entry changes, so each pair meets the same test it fails here.  Stop now;
    // the diagonal and the recorded rotations are those the rotation-free sweep would leave (it only
    // clears the off-diagonal entries below the test, which nothing reads).
    if (sweep + 1 < max_sweeps) {
      int any = 0;
#pragma unroll
      for (int u = 0; u < JB_MAX; ++u) {
        if (bij[u] < 0) continue;
        const int i = bij[u] & 0xffff, j = bij[u] >> 16, b = t + u * JR_NT;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (i == j && e != 1) continue;   // a diagonal block holds one off-diagonal entry, in element 1
          const int x = 2 * i + (e >> 1), y = 2 * j + (e & 1);
          const double axx = a[((x & 1) ? 3 * nblk : 0) + jblk(x >> 1, x >> 1)];
          const double ayy = a[((y & 1) ? 3 * nblk : 0) + jblk(y >> 1, y >> 1)];
          any |= jrot_needed(axx, a[e * nblk + b], ayy, abs_floor) ? 1 : 0;
        }
      }
      if (!__syncthreads_or(any)) { ++sweep; break; }
    }
  }
  // whole sweeps ran: every index is back at its round-0 position
  for (int i = t; i < m; i += JR_NT) {
    int pos;
    if (i == 0) pos = 0;
    else if (i == M) pos = 1;
    else if (i < P) pos = 2 * i;
    else pos = 2 * (M - i) + 1;
    diag[i] = a[((pos & 1) ? 3 * nblk : 0) + jblk(pos >> 1, pos >> 1)];
  }
  if (t == 0) {
    *rounds_out = round;
    if (sweeps_out) *sweeps_out = sweep;
  }
}

// Replays the recorded rotations on rows of V (V = product of all rotations, starting from the
// identity).  Rows are independent: each wavefront owns one row in LDS and walks through the rounds
// on its own (the rotations of one round touch disjoint column pairs, and a wavefront's LDS accesses
// complete in order), so workgroup barriers are only needed to stage rotation parameters.
constexpr int JA_WAVES = 4;   // rows (wavefronts) per workgroup
constexpr int JCHUNK = 16;    // rounds of rotation parameters staged in LDS at once
__global__ __launch_bounds__(64 * JA_WAVES) void jacobi_apply_kernel(const double* __restrict__ rot,
                                                                    const int* __restrict__ rounds_ptr, int m,
                                                                    double* __restrict__ Vrows,
                                                                    const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double rows[JA_WAVES][JACOBI_MAX];
  __shared__ double2 prm[JCHUNK * (JACOBI_MAX / 2)];
  const int mm = m + (m & 1);
  const int P = mm / 2;
  const int M = mm - 1;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int row = blockIdx.x * JA_WAVES + wv;
  const bool live = row < m;
  double* R = rows[wv];
  for (int c = lane; c < mm; c += 64) R[c] = (c == row) ? 1.0 : 0.0;
  const int Rn = *rounds_ptr;
  const double2* src = reinterpret_cast<const double2*>(rot);
  for (int base = 0; base < Rn; base += JCHUNK) {
    const int nr = min(JCHUNK, Rn - base);
    __syncthreads();   // the previous chunk is consumed
    for (int e = t; e < nr * P; e += 64 * JA_WAVES) prm[e] = src[(size_t)base * P + e];
    __syncthreads();
    if (live) {
      int r_in = base % M;
      // a lane turns slots lane and lane + 64 (P <= 96); the two pairs are disjoint, so both are
      // read before either is written
      const bool two = lane + 64 < P;
      for (int rr = 0; rr < nr; ++rr) {
        const double2 cs0 = prm[rr * P + lane];
        const double2 cs1 = two ? prm[rr * P + lane + 64] : make_double2(1.0, 0.0);
        int p0 = 0, q0 = 0, p1 = 0, q1 = 0;   // lanes without a slot read entry 0 and write nothing
        if (lane < P) rr_pair(r_in, lane, M, p0, q0);
        if (two) rr_pair(r_in, lane + 64, M, p1, q1);
        const bool act0 = lane < P && cs0.y != 0.0, act1 = cs1.y != 0.0;
        const double v0p = R[p0], v0q = R[q0], v1p = R[p1], v1q = R[q1];
        if (act0) {
          R[p0] = cs0.x * v0p - cs0.y * v0q;
          R[q0] = cs0.y * v0p + cs0.x * v0q;
        }
        if (act1) {
          R[p1] = cs1.x * v1p - cs1.y * v1q;
          R[q1] = cs1.y * v1p + cs1.x * v1q;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (++r_in == M) r_in = 0;
      }
    }
  }
  if (live)
    for (int c = lane; c < m; c += 64) Vrows[row + (size_t)m * c] = R[c];   // column-major V
}

// sort by |lambda| descending (stable), sign convention: largest |component| of each vector positive
// (the first such component when several tie).  One workgroup per (unsorted) eigenpair.
constexpr int JS_NT = 256;
__global__ __launch_bounds__(JS_NT) void jacobi_sort_kernel(const double* __restrict__ diag,
                                                            const double* __restrict__ Vu, int m,
                                                            double* __restrict__ evals, double* __restrict__ V,
                                                            const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ int cnt_w[JS_NT / 64], row_w[JS_NT / 64];
  __shared__ double best_w[JS_NT / 64];
  const int i = blockIdx.x, t = threadIdx.x;
  const double li = diag[i], ai = fabs(li);
  const double* vi = Vu + (long long)i * m;
  int cnt = 0;
  double best = -1.0;
  int brow = m;
  for (int j = t; j < m; j += JS_NT) {
    const double aj = fabs(diag[j]);
    cnt += (aj > ai) || (aj == ai && j < i);
    const double v = fabs(vi[j]);
    if (v > best) { best = v; brow = j; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cnt += __shfl_down(cnt, o);
    const double ob = __shfl_down(best, o);
    const int orow = __shfl_down(brow, o);
    if (ob > best || (ob == best && orow < brow)) { best = ob; brow = orow; }
  }
  if ((t & 63) == 0) { cnt_w[t >> 6] = cnt; best_w[t >> 6] = best; row_w[t >> 6] = brow; }
  __syncthreads();
  int rank = 0;
  double bb = -1.0;
  int br = m;
  for (int w = 0; w < JS_NT / 64; ++w) {
    rank += cnt_w[w];
    if (best_w[w] > bb || (best_w[w] == bb && row_w[w] < br)) { bb = best_w[w]; br = row_w[w]; }
  }
  const double sgn = (br < m && vi[br] < 0) ? -1.0 : 1.0;
  if (t == 0) evals[rank] = li;
  for (int row = t; row < m; row += JS_NT) V[row + (long long)rank * m] = sgn * vi[row];
}

// Cholesky (with diagonal shift) + triangular inverse of an m x m SPD matrix, m <= 192,
// in one workgroup of 16 x 16 threads (one wavefront per SIMD: the per-step costs that every
// thread pays, the pivot's reciprocal square root and the scaling of its column entries, are
// then paid by 4 wavefronts instead of 16; 158 -> 140 us at m = 192).  Thread (ty, tx) holds
// the 12 x 12 elements (ty + 16a, tx + 16b) in registers (cyclic distribution, so every step
// keeps all threads busy).  One sweep of m elimination steps forms both L and B = L^-1: step jj
// broadcasts column jj of the Schur complement and row jj of B (both through LDS, one
// barrier per step, double-buffered); every thread then updates its part of the trailing
// Schur complement (A -= l l^T) and of the rows of B below jj (B[i,:] -= L[i,jj] B[jj,:],
// forward substitution for L B = I), and the owners finish column jj of L and row jj of B.
// Only L's current column is ever needed, so finished columns of L are dropped, and the
// rows of B in a 16-row block are final when the block's steps are done: they go straight
// to the output.  The steps run in twelve 16-column blocks whose block index is a template
// parameter, so every register access is static and blocks that a step cannot touch
// (finished columns, the upper triangle) are not even emitted.
// Output Rinv = L^-T (upper triangular, column-major, zeros below the diagonal) so
// that X = Y Rinv; piv (nullable) receives diag(L).
constexpr int CH_G = 16;                  // thread grid edge
constexpr int CH_R = JACOBI_MAX / CH_G;   // elements per thread per dimension
static_assert(CH_R * CH_G == JACOBI_MAX, "Cholesky tiling");

struct CholShared {
  double* cb;     // 2 x JACOBI_MAX broadcast buffers: column jj of the Schur complement
  double* rb;     // 2 x JACOBI_MAX broadcast buffers: row jj of B before its scaling
};

// the first non-positive pivot (1-based, 0: none) and the smallest pivot so far, kept by thread 0 in
// registers (every thread sees the same pivots)
struct CholState {
  int bad = 0;
  double pmin = HUGE_VAL;
};

// 1 / sqrt(d) for a positive normal d: the hardware estimate refined by two Newton steps
__device__ __forceinline__ double rsqrt_refined(double d) {
  double y = __builtin_amdgcn_rsq(d);
  y = y * (1.5 - 0.5 * d * y * y);
  y = y * (1.5 - 0.5 * d * y * y);
  return y;
}

// elimination steps jj = 16 BJ .. min(16 BJ + 15, m - 1).  The owner column broadcasts the whole
// column jj (zeros above the diagonal and in the padding), so the update needs no bounds tests: its
// only stray products land in row jj above the diagonal (the unused upper triangle) and in column
// jj, which the owner rewrites.  Likewise row jj of B is zero right of its diagonal, its stray
// product lands in row jj itself (rewritten by its owner), and rows above jj meet a zero L[i,jj].
// After the block's steps its rows of B are final and are written to Rinv.
template <int BJ>
__device__ __forceinline__ void chol_block(double (&A)[CH_R][CH_R], double (&B)[CH_R][CH_R], int m,
                                           const CholShared& sh, CholState& st, int tx, int ty, int t,
                                           double* __restrict__ piv, double* __restrict__ Rinv) {
  const int j0 = BJ * CH_G, j1 = min(m, j0 + CH_G);
  for (int jj = j0; jj < j1; ++jj) {
    double* col = sh.cb + (jj & 1) * JACOBI_MAX;
    double* row = sh.rb + (jj & 1) * JACOBI_MAX;
    const int tj = jj - j0;
    if (tx == tj) {
#pragma unroll
      for (int a = BJ; a < CH_R; ++a) {
        const int i = ty + CH_G * a;
        col[i] = (i >= jj && i < m) ? A[a][BJ] : 0.0;
      }
    }
    if (ty == tj) {
#pragma unroll
      for (int b = 0; b <= BJ; ++b) row[tx + CH_G * b] = B[BJ][b];
    }
    __syncthreads();
    // every LDS read of the step first, so that their latency overlaps the pivot's reciprocal square root;
    // no branch before the products (the pivot bookkeeping is thread 0's, in registers)
    double d = col[jj];
    double li[CH_R], lj[CH_R], xr[CH_R];
#pragma unroll
    for (int a = BJ; a < CH_R; ++a) {
      li[a] = col[ty + CH_G * a];
      lj[a] = col[tx + CH_G * a];
    }
#pragma unroll
    for (int b = 0; b <= BJ; ++b) xr[b] = row[tx + CH_G * b];
    const bool ok = d > 0.0;
    st.bad = (!ok && st.bad == 0) ? jj + 1 : st.bad;
    d = ok ? d : 1e-300;
    const double isd = rsqrt_refined(d);
    const double sd = d * isd;
    st.pmin = fmin(st.pmin, sd);
#pragma unroll
    for (int a = BJ; a < CH_R; ++a) {
      li[a] *= isd;
      lj[a] *= isd;
    }
#pragma unroll
    for (int b = 0; b <= BJ; ++b) xr[b] *= isd;   // row jj of B, final
    if (t == 0 && piv) piv[jj] = sd;
    // trailing lower blocks BJ <= b <= a of the Schur complement
#pragma unroll
    for (int a = BJ; a < CH_R; ++a)
#pragma unroll
      for (int b = BJ; b <= a; ++b) A[a][b] -= li[a] * lj[b];
    // rows of B below jj (blocks a >= BJ, columns b <= BJ)
#pragma unroll
    for (int a = BJ; a < CH_R; ++a)
#pragma unroll
      for (int b = 0; b <= BJ; ++b) B[a][b] -= li[a] * xr[b];
    if (tx == tj) {
#pragma unroll
      for (int a = BJ; a < CH_R; ++a) {
        const int i = ty + CH_G * a;
        if (i >= jj) A[a][BJ] = (i == jj) ? sd : li[a];
      }
    }
    if (ty == tj) {
#pragma unroll
      for (int b = 0; b <= BJ; ++b) B[BJ][b] = xr[b];
    }
  }
  // rows i = ty + 16 BJ of B are final: Rinv(j, i) = B(i, j), zero above B's diagonal
  const int i = ty + CH_G * BJ;
  if (i < m) {
#pragma unroll
    for (int b = 0; b < CH_R; ++b) {
      const int j = tx + CH_G * b;
      if (j < m) Rinv[j + (size_t)m * i] = (b <= BJ && j <= i) ? B[BJ][b <= BJ ? b : 0] : 0.0;
    }
  }
}

// the column blocks BJ, BJ + 1, ..., CH_R - 1 in turn
template <int BJ>
__device__ __forceinline__ void chol_blocks(double (&A)[CH_R][CH_R], double (&B)[CH_R][CH_R], int m,
                                            const CholShared& sh, CholState& st, int tx, int ty, int t,
                                            double* __restrict__ piv, double* __restrict__ Rinv) {
  chol_block<BJ>(A, B, m, sh, st, tx, ty, t, piv, Rinv);
  if constexpr (BJ + 1 < CH_R) chol_blocks<BJ + 1>(A, B, m, sh, st, tx, ty, t, piv, Rinv);
}

__global__ __launch_bounds__(CH_G * CH_G) void chol_inv_kernel(const double* __restrict__ G, int m,
                                                               double shift, double* __restrict__ Rinv,
                                                               int* info, double* __restrict__ piv,
                                                               double* __restrict__ pmin_out) {
  __shared__ double bufs[4 * JACOBI_MAX];
  CholShared sh{bufs, bufs + 2 * JACOBI_MAX};
  CholState st;
  const int t = threadIdx.x, tx = t % CH_G, ty = t / CH_G;
  double A[CH_R][CH_R], B[CH_R][CH_R];
#pragma unroll
  for (int a = 0; a < CH_R; ++a) {
#pragma unroll
    for (int b = 0; b < CH_R; ++b) {
      const int i = ty + CH_G * a, j = tx + CH_G * b;
      const double g = G[min(j, m - 1) + (size_t)m * min(i, m - 1)];   // branch-free, clamped address
      // padding beyond m: identity, never coupled to the rest
      A[a][b] = (i < m && j < m) ? g + (i == j ? shift : 0.0) : (i == j ? 1.0 : 0.0);
      B[a][b] = (i == j) ? 1.0 : 0.0;
    }
  }
  chol_blocks<0>(A, B, m, sh, st, tx, ty, t, piv, Rinv);
  if (t == 0) {
    if (info) *info = st.bad;
    if (pmin_out) *pmin_out = st.pmin;
  }
}


// Rayleigh-Ritz of a nearly diagonal symmetric H (m <= 192, one workgroup) for its first nh
// ("kept") columns.  Off-diagonal entries at or below the rounding level of H (4.5e-16 |H|_F, where
// the Jacobi sweeps stop as well) count as zero.
//  1. Pairs that involve a kept column and are strongly coupled, |H_ij| > thresh |H_ii - H_jj|
//     (near-degenerate pairs, which perturbation theory cannot treat), are removed by Jacobi
//     rotations on the packed triangle in LDS: each pass rotates a greedy set of disjoint strong
//     pairs, until none is left (at most RS_PASSES passes, RS_MAXROT rotations, RS_LIST strong pairs
//     per pass).  A rotation keeps the larger diagonal entry at the smaller index, so the kept
//     positions keep the larger values.
//  2. The rotated matrix, diagonal d plus weak couplings E (restricted to pairs with a kept column),
//     has the eigenpairs, to second order in E,
//        theta_j = d_j + sum_{i != j} E_ij^2 / (d_j - d_i),
//        v_j = (e_j + sum_{i != j} e_i E_ij / (d_j - d_i)) / norm,
//     and those of H are the rotations of step 1 applied to v_j.  Pairs among the further columns
//     (i, j >= nh) are left unrotated: those vectors are not converged and only widen the search
//     space, which any mixing among them leaves unchanged -- provided the further block cannot hold
//     a direction that belongs among the kept ones, which its Gershgorin bound (below the smallest
//     kept |d_i| by 0.1%) rules out.
// The pairs are written sorted by |theta| descending (stable) and *flag = 0.  When a limit of step 1
// is hit or the Gershgorin test fails, *flag > 0 names the reason (2: more strong pairs than the list
// holds, 3: passes, 4: rotation budget, 5: Gershgorin) and nothing else is written (the caller falls back
// to Jacobi).  The vectors are orthonormal to O((E/gap)^2); the subspace iteration's residual test
// sees any remainder.
constexpr int RS_LIST = 256;
constexpr int RS_PASSES = 12;
constexpr int RS_MAXROT = JACOBI_MAX;

__global__ __launch_bounds__(EIG_NT) void ritz_select_kernel(const double* __restrict__ H, int ldh, int m,
                                                             int nh, double thresh, double* __restrict__ evals,
                                                             double* __restrict__ V, int* __restrict__ flag) {
  extern __shared__ double smem[];
  double* a = smem;   // packed upper triangle, pk(i, j)
  __shared__ double dg[JACOBI_MAX], th[JACOBI_MAX], nr[JACOBI_MAX];
  __shared__ int rk[JACOBI_MAX], selid[JACOBI_MAX];
  __shared__ int lst[RS_LIST], lsts[RS_LIST];
  __shared__ double rc[RS_MAXROT], rsn[RS_MAXROT];
  __shared__ int rpq[RS_MAXROT];
  __shared__ double red[EIG_NT / 64], red2[EIG_NT / 64];
  __shared__ int cnt_s, nsel_s, nrot_s, bad_s;
  __shared__ double fro_s;
  const int t = threadIdx.x;
  const int ntri = m * (m + 1) / 2;
  double part = 0.0;
  for (int b = t; b < ntri; b += EIG_NT) {
    int i, j;
    decode_tri(b, i, j);
    const double v = 0.5 * (H[i + (long long)ldh * j] + H[j + (long long)ldh * i]);
    a[b] = v;
    part += (i == j ? 1.0 : 2.0) * v * v;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o);
  if ((t & 63) == 0) red[t >> 6] = part;
  if (t == 0) { nrot_s = 0; bad_s = 0; }
  __syncthreads();
  if (t == 0) {
    double acc = 0.0;
    for (int w = 0; w < EIG_NT / 64; ++w) acc += red[w];
    fro_s = sqrt(acc);
  }
  __syncthreads();
  const double floor_abs = 4.5e-16 * fro_s + 1e-300;

  // ---- 1. rotate the strongly coupled pairs away ----
  for (int pass = 0;; ++pass) {
    for (int i = t; i < m; i += EIG_NT) { dg[i] = a[pk(i, i)]; selid[i] = -1; }
    if (t == 0) cnt_s = 0;
    __syncthreads();
    for (int b = t; b < ntri; b += EIG_NT) {
      int i, j;
      decode_tri(b, i, j);
      if (i == j || i >= nh) continue;   // i < j: a pair with a kept column has i < nh
      const double e = fabs(a[b]);
      if (e <= floor_abs || e <= thresh * fabs(dg[i] - dg[j])) continue;
      const int slot = atomicAdd(&cnt_s, 1);
      if (slot < RS_LIST) lst[slot] = i | (j << 16);
    }
    __syncthreads();
    const int cnt = cnt_s;
    if (cnt == 0) break;
    if (cnt > RS_LIST || pass == RS_PASSES) {
      if (t == 0) bad_s = cnt > RS_LIST ? 2 : 3;
      break;
    }
    // the list fills in arrival order: sort it (keys are distinct) so that the selection, and with it
    // the result, does not depend on thread timing
    for (int u = t; u < cnt; u += EIG_NT) {
      const int key = lst[u];
      int r = 0;
      for (int w = 0; w < cnt; ++w) r += lst[w] < key;
      lsts[r] = key;
    }
    __syncthreads();
    if (t == 0) {   // greedy set of disjoint pairs, within the rotation budget
      const int r0 = nrot_s;
      int ns = 0;
      for (int u = 0; u < cnt && r0 + ns < RS_MAXROT; ++u) {
        const int i = lsts[u] & 0xffff, j = lsts[u] >> 16;
        if (selid[i] < 0 && selid[j] < 0) {
          selid[i] = ns;
          selid[j] = ns;
          rpq[r0 + ns] = i | (j << 16);
          ++ns;
        }
      }
      nsel_s = ns;
      if (ns == 0) bad_s = 4;
    }
    __syncthreads();
    if (bad_s) break;
    const int ns = nsel_s, r0 = nrot_s;
    if (t < ns) {
      const int p = rpq[r0 + t] & 0xffff, q = rpq[r0 + t] >> 16;
      const double app = dg[p], aqq = dg[q], apq = a[pk(p, q)];
      const double theta = (aqq - app) / (2.0 * apq);
      const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      double c = 1.0 / sqrt(tt * tt + 1.0), sn = tt * c;
      if (fabs(app - tt * apq) < fabs(aqq + tt * apq)) {   // swap the two new directions
        const double c2 = sn;
        sn = -c;
        c = c2;
      }
      rc[r0 + t] = c;
      rsn[r0 + t] = sn;
    }
    __syncthreads();
    // A <- J^T A J.  Work item (u, k): selected pair u = (p, q) with index k.  k outside every selected
    // pair: the two entries (p, k), (q, k) turn by pair u alone; k = the first index of a selected pair
    // v > u: the 2x2 block of pairs u and v turns on both sides; k = p: the diagonal block of pair u.
    for (int e = t; e < ns * m; e += EIG_NT) {
      const int u = e / m, k = e % m;
      const int p = rpq[r0 + u] & 0xffff, q = rpq[r0 + u] >> 16;
      const double ci = rc[r0 + u], si = rsn[r0 + u];
      const int v = selid[k];
      if (v < 0) {
        const double xp = a[pku(p, k)], xq = a[pku(q, k)];
        a[pku(p, k)] = ci * xp - si * xq;
        a[pku(q, k)] = si * xp + ci * xq;
      } else if (v == u) {
        if (k != p) continue;
        const double app = a[pk(p, p)], aqq = a[pk(q, q)], apq = a[pk(p, q)];
        const double c2 = ci * ci, s2 = si * si, cs = ci * si;
        a[pk(p, p)] = c2 * app - 2.0 * cs * apq + s2 * aqq;
        a[pk(q, q)] = s2 * app + 2.0 * cs * apq + c2 * aqq;
        a[pk(p, q)] = 0.0;
      } else {
        const int r = rpq[r0 + v] & 0xffff, s = rpq[r0 + v] >> 16;
        if (v < u || k != r) continue;
        const double cj = rc[r0 + v], sj = rsn[r0 + v];
        const double xpr = a[pku(p, r)], xps = a[pku(p, s)], xqr = a[pku(q, r)], xqs = a[pku(q, s)];
        const double ypr = ci * xpr - si * xqr, yps = ci * xps - si * xqs;
        const double yqr = si * xpr + ci * xqr, yqs = si * xps + ci * xqs;
        a[pku(p, r)] = cj * ypr - sj * yps;
        a[pku(p, s)] = sj * ypr + cj * yps;
        a[pku(q, r)] = cj * yqr - sj * yqs;
        a[pku(q, s)] = sj * yqr + cj * yqs;
      }
    }
    __syncthreads();
    if (t == 0) nrot_s = r0 + ns;
    __syncthreads();
  }
  __syncthreads();
  if (bad_s) {
    if (t == 0) *flag = bad_s;
    return;
  }
  const int nrot = nrot_s;

  // ---- 2. the further block must not hide a kept direction (Gershgorin) ----
  double gmax = 0.0, hmin = HUGE_VAL;
  for (int i = nh + t; i < m; i += EIG_NT) {
    double rs = 0.0;
    for (int j = nh; j < m; ++j) rs += fabs(a[pku(i, j)]);
    gmax = fmax(gmax, rs);
  }
  for (int i = t; i < nh; i += EIG_NT) hmin = fmin(hmin, fabs(dg[i]));
  for (int o = 32; o > 0; o >>= 1) {
    gmax = fmax(gmax, __shfl_down(gmax, o));
    hmin = fmin(hmin, __shfl_down(hmin, o));
  }
  if ((t & 63) == 0) { red[t >> 6] = gmax; red2[t >> 6] = hmin; }
  __syncthreads();
  if (t == 0) {
    double g = 0.0, hm = HUGE_VAL;
    for (int w = 0; w < EIG_NT / 64; ++w) { g = fmax(g, red[w]); hm = fmin(hm, red2[w]); }
    bad_s = !(g < (1.0 - 1e-3) * hm) ? 5 : 0;
    if (bad_s) *flag = bad_s;
  }
  __syncthreads();
  if (bad_s) return;

  // ---- 3. second-order eigenpairs of the rotated matrix, rotated back ----
  auto E = [&](int i, int j) -> double {
    if (i >= nh && j >= nh) return 0.0;
    const double e = a[pku(i, j)];
    return fabs(e) <= floor_abs ? 0.0 : e;   // a nonzero E has |E| <= thresh |d_i - d_j|: the gap is > 0
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
  // one thread per output column: it writes the column and turns it back through the rotations
  // (last first: H's eigenvectors are J_1 ... J_n v), touching only its own column
  for (int j = t; j < m; j += EIG_NT) {
    double* v = V + (long long)m * rk[j];
    for (int i = 0; i < m; ++i) {
      const double eij = (i == j) ? 0.0 : E(i, j);
      v[i] = (i == j ? 1.0 : (eij == 0.0 ? 0.0 : eij / (dg[j] - dg[i]))) * nr[j];
    }
    for (int u = nrot - 1; u >= 0; --u) {
      const int p = rpq[u] & 0xffff, q = rpq[u] >> 16;
      const double c = rc[u], sn = rsn[u];
      const double xp = v[p], xq = v[q];
      v[p] = c * xp + sn * xq;
      v[q] = -sn * xp + c * xq;
    }
  }
  if (t == 0) *flag = 0;
}
}  // namespace

void sort_eigenpairs(const double* lam, const double* Vu, int m, double* evals, double* V, const int* gate,
                     hipStream_t s) {
  hipLaunchKernelGGL(jacobi_sort_kernel, dim3(m), dim3(JS_NT), 0, s, lam, Vu, m, evals, V, gate);
  TNAD_HIP_TRY(hipGetLastError());
}

static size_t jacobi_lds(int m) {   // four planes of P (P + 1) / 2 blocks, P = ceil(m / 2)
  const int P = (m + 1) / 2;
  return (size_t)(4 * (P * (P + 1) / 2)) * sizeof(double);
}

size_t jacobi_work_doubles(int m, int max_sweeps) {
  const int mm = m + (m & 1);
  const size_t P = mm / 2;
  // rotation record + unsorted V + diagonal + round counter
  return (size_t)max_sweeps * (mm - 1) * P * 2 + (size_t)m * m + m + 8;
}

void jacobi_eig_ws(const double* H, int ldh, int m, double* evals, double* V, double* Vwork, int max_sweeps,
                   int* sweeps_out, hipStream_t s, const int* gate) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "jacobi_eig: size out of range");
  if (max_sweeps < 1 || max_sweeps > JR_MAX_SWEEPS) throw Error(TNAD_ERR_ARG, "jacobi_eig: sweep limit out of range");
  const int mm = m + (m & 1);
  const size_t P = mm / 2;
  double* rot = Vwork;
  double* Vu = rot + (size_t)max_sweeps * (mm - 1) * P * 2;
  double* diag = Vu + (size_t)m * m;
  int* rounds = reinterpret_cast<int*>(diag + m);
  size_t lds = jacobi_lds(m);
  const bool next = (long long)P * (P + 1) / 2 > (long long)(JB_MAX - 1) * JR_NT;
  const void* kern = next ? (const void*)jacobi_rot_kernel<true> : (const void*)jacobi_rot_kernel<false>;
  TNAD_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (next)
    hipLaunchKernelGGL(jacobi_rot_kernel<true>, dim3(1), dim3(JR_NT), lds, s, H, ldh, m, rot, max_sweeps, diag,
                       rounds, sweeps_out, gate);
  else
    hipLaunchKernelGGL(jacobi_rot_kernel<false>, dim3(1), dim3(JR_NT), lds, s, H, ldh, m, rot, max_sweeps, diag,
                       rounds, sweeps_out, gate);
  TNAD_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(jacobi_apply_kernel, dim3((m + JA_WAVES - 1) / JA_WAVES), dim3(64 * JA_WAVES), 0, s, rot,
                     rounds, m, Vu, gate);
  TNAD_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(jacobi_sort_kernel, dim3(m), dim3(JS_NT), 0, s, diag, Vu, m, evals, V, gate);
  TNAD_HIP_TRY(hipGetLastError());
}

float jacobi_rot_time(const double* H, int m, int max_sweeps, int reps, double* Vwork, hipStream_t s,
                      int* rounds_dev) {
  if (m < 2 || m > JACOBI_MAX || reps < 1) throw Error(TNAD_ERR_ARG, "jacobi_rot_time: bad sizes");
  if (max_sweeps < 1 || max_sweeps > JR_MAX_SWEEPS) throw Error(TNAD_ERR_ARG, "jacobi_rot_time: sweep limit out of range");
  const int mm = m + (m & 1);
  const size_t P = mm / 2;
  double* rot = Vwork;
  double* diag = rot + (size_t)max_sweeps * (mm - 1) * P * 2 + (size_t)m * m;
  const size_t lds = jacobi_lds(m);
  const bool next = (long long)P * (P + 1) / 2 > (long long)(JB_MAX - 1) * JR_NT;
  const void* kern = next ? (const void*)jacobi_rot_kernel<true> : (const void*)jacobi_rot_kernel<false>;
  TNAD_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  auto launch = [&] {
    if (next)
      hipLaunchKernelGGL(jacobi_rot_kernel<true>, dim3(1), dim3(JR_NT), lds, s, H, m, m, rot, max_sweeps, diag,
                         rounds_dev, (int*)nullptr, (const int*)nullptr);
    else
      hipLaunchKernelGGL(jacobi_rot_kernel<false>, dim3(1), dim3(JR_NT), lds, s, H, m, m, rot, max_sweeps, diag,
                         rounds_dev, (int*)nullptr, (const int*)nullptr);
    TNAD_HIP_TRY(hipGetLastError());
  };
  launch();   // warm-up
  hipEvent_t e0, e1;
  TNAD_HIP_TRY(hipEventCreate(&e0));
  TNAD_HIP_TRY(hipEventCreate(&e1));
  TNAD_HIP_TRY(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) launch();
  TNAD_HIP_TRY(hipEventRecord(e1, s));
  TNAD_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  TNAD_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

void ritz_select(const double* H, int ldh, int m, int nh, double thresh, double* evals, double* V, int* flag,
                 hipStream_t s) {
  if (m < 1 || m > JACOBI_MAX || nh < 1 || nh > m) throw Error(TNAD_ERR_ARG, "ritz_select: size out of range");
  const size_t lds = (size_t)m * (m + 1) / 2 * sizeof(double);
  TNAD_HIP_TRY(hipFuncSetAttribute((const void*)ritz_select_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
  hipLaunchKernelGGL(ritz_select_kernel, dim3(1), dim3(EIG_NT), lds, s, H, ldh, m, nh, thresh, evals, V, flag);
  TNAD_HIP_TRY(hipGetLastError());
}

void chol_inv(const double* G, int m, double shift, double* Rinv, int* info, double* piv, hipStream_t s,
              double* pmin) {
  if (m < 1 || m > JACOBI_MAX) throw Error(TNAD_ERR_ARG, "chol_inv: size out of range");
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(CH_G * CH_G), 0, s, G, m, shift, Rinv, info, piv, pmin);
  TNAD_HIP_TRY(hipGetLastError());
}

}  // namespace tnad
