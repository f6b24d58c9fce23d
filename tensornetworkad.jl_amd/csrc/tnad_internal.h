// Internal declarations shared by the HIP translation units of libtnad.
// Layout convention everywhere: column-major (first index fastest), the
// memory order of Julia arrays, so buffers cross the C ABI without transposes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>
#include <map>

#include "tnad_limits.h"
#include "tnad.h"

#define TNAD_HIP_TRY(expr)                                                          \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      throw tnad::Error(TNAD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    }                                                                               \
  } while (0)

namespace tnad {

struct Error {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// ---------------------------------------------------------------------------
// Device buffer with RAII; allocations go through the context's cache.
// ---------------------------------------------------------------------------
struct Ctx;

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  Ctx* owner = nullptr;
  DevBuf() = default;
  DevBuf(Ctx* c, size_t b);
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
  DevBuf& operator=(DevBuf&& o) noexcept;
  ~DevBuf();
  template <typename T> T* as() const { return reinterpret_cast<T*>(ptr); }
};

// ---------------------------------------------------------------------------
// Generic strided contraction  C[m,n] = alpha * sum_k A[m,k] B[k,n] + beta * C[m,n]
// Each of m, n, k is a pair of sub-indices (x0 fastest) with arbitrary element
// strides, so every einsum of the hot path is one launch with no permute pass.
// ---------------------------------------------------------------------------
struct TcDesc {
  int M0 = 1, M1 = 1, N0 = 1, N1 = 1, K0 = 1, K1 = 1;
  long long am0 = 0, am1 = 0, ak0 = 0, ak1 = 0;
  long long bk0 = 0, bk1 = 0, bn0 = 0, bn1 = 0;
  long long cm0 = 0, cm1 = 0, cn0 = 0, cn1 = 0;
  int batch = 1;
  long long abat = 0, bbat = 0, cbat = 0;
  double alpha = 1.0, beta = 0.0;
  int kchunk = 0;   // set by tc_launch: K elements per split-K slice (0: no split)
};

// Plain column-major GEMM helper: C(MxN) = alpha op(A) op(B) + beta C
TcDesc gemm_desc(char ta, char tb, int M, int N, int K, long long lda, long long ldb, long long ldc,
                 double alpha = 1.0, double beta = 0.0);

enum class Prec { F64 = TNAD_PREC_FP64, F32 = TNAD_PREC_FP32, BF16 = TNAD_PREC_BF16 };

struct Ctx;
// A, B, C in double (F64) or float (F32 / BF16: BF16 rounds operands on the way into LDS).
// Every operand's full strided extent is checked against the context's live device
// blocks before the launch (a size or precision mismatch throws instead of faulting).
void tc_launch(Ctx* c, Prec p, const TcDesc& d, const void* A, const void* B, void* C, hipStream_t s);

// ---------------------------------------------------------------------------
// Small dense kernels (eig.hip)
// ---------------------------------------------------------------------------
// JACOBI_MAX (192): the packed FP64 upper triangle of 192x192 fits in LDS (tnad_limits.h)
inline void enforce(const Limit& l) {
  if (l) throw Error(l.code, l.msg);
}

// Full symmetric eigendecomposition of an m x m matrix (m <= JACOBI_MAX).
// Output eigenpairs sorted by |lambda| descending; V column-major (ld = m).
// Vwork: jacobi_work_doubles(m, max_sweeps) doubles of scratch (rotation record
// and unsorted eigenvectors).  sweeps_out: device int (nullable).  gate (nullable device int): when
// it holds 0 at run time the kernels return at once and evals / V keep what they hold (a fallback
// that the device decides, without a host round trip).
size_t jacobi_work_doubles(int m, int max_sweeps);
// Average device time (ms) of one rotation-kernel launch of the Jacobi solver on H (m x m, ld m; at most
// max_sweeps sweeps) over reps launches after one warm-up, HIP events on stream s; *rounds_dev (device int)
// receives the rounds each launch ran.  Vwork as for jacobi_eig_ws.  For roofline measurement.
float jacobi_rot_time(const double* H, int m, int max_sweeps, int reps, double* Vwork, hipStream_t s,
                      int* rounds_dev);
void jacobi_eig_ws(const double* H, int ldh, int m, double* evals, double* V, double* Vwork,
                   int max_sweeps, int* sweeps_out, hipStream_t s, const int* gate = nullptr);

// Full symmetric eigendecomposition of an m x m matrix (m <= JACOBI_MAX, the upper triangle of H is read)
// by Householder tridiagonalisation, divide and conquer and back-transformation (dc.hip); output as
// jacobi_eig_ws (sorted by |lambda| descending, sign convention).  work: syev_dc_work_doubles(m) doubles.
size_t syev_dc_work_doubles(int m);
void syev_dc(const double* H, int ldh, int m, double* evals, double* V, double* work, hipStream_t s,
             const int* gate = nullptr);
// Average device times (ms) of syev_dc's stages on H (m x m, ld m; 3 <= m <= 192) over reps runs after one
// warm-up, HIP events on stream s: ms[0] tridiagonalisation, ms[1] divide and conquer, ms[2]
// back-transformation and sort.  work: syev_dc_work_doubles(m) + m * m + m doubles.  For roofline measurement.
void syev_dc_time(const double* H, int m, int reps, double* work, hipStream_t s, float ms[3]);
// eigenpairs (lam[i], column i of Vu, m x m) sorted by |lambda| descending into evals / V with the sign
// convention of jacobi_eig_ws; skipped when *gate == 0
void sort_eigenpairs(const double* lam, const double* Vu, int m, double* evals, double* V, const int* gate,
                     hipStream_t s);

// Eigenpairs of a nearly diagonal symmetric m x m H for its first nh columns (pairs among the
// others are not rotated), sorted by |lambda| descending: pairs with min(i, j) < nh coupled more
// strongly than |H_ij| > thresh |H_ii - H_jj| are rotated away by Jacobi rotations (a few passes),
// the rest follows by second-order perturbation theory.  Writes *flag = 0 (device int) on success;
// *flag = 1 leaves evals / V untouched (too many strong pairs, or the further block could hide a
// kept direction).
void ritz_select(const double* H, int ldh, int m, int nh, double thresh, double* evals, double* V, int* flag,
                 hipStream_t s);

// Cholesky of the m x m SPD matrix G (+ shift on the diagonal), returns Linv^T
// (upper triangular, column-major) so that X = Y * Rinv with Rinv = L^-T.
// piv (nullable, m doubles) receives the Cholesky pivots diag(L); pmin (nullable, device double) the
// smallest of them.
void chol_inv(const double* G, int m, double shift, double* Rinv, int* info, double* piv, hipStream_t s,
              double* pmin = nullptr);

// ---------------------------------------------------------------------------
// Elementwise / reduction kernels (elem.hip)
// ---------------------------------------------------------------------------
void k_sym_add(const double* X, int n, double* Y, hipStream_t s);              // Y = X + X^T
template <typename T> void k_cast(const T* src, double* dst, size_t n, hipStream_t s);
template <typename T> void k_cast_from(const double* src, T* dst, size_t n, hipStream_t s);
void k_fill(double* x, size_t n, double v, hipStream_t s);
void k_copy(const double* x, double* y, size_t n, hipStream_t s);
void k_axpby(double a, const double* x, double b, double* y, size_t n, hipStream_t s); // y = a x + b y
void k_scale_by_dev(double* x, size_t n, const double* nrm, hipStream_t s);   // x /= *nrm
void k_sumsq(const double* x, size_t n, double* out, hipStream_t s);           // *out = sum x^2 (atomic-free 2 pass)
void k_dot(const double* x, const double* y, size_t n, double* out, hipStream_t s);
void k_edge_sym(const double* T1, int chi, int D, double* T2, hipStream_t s);  // T2 = T1 + T1^{kji}
void k_col_norms(const double* X, int n, int k, double* nrm, hipStream_t s);
void k_scale_cols(double* X, int n, int k, const double* scale, int invert, hipStream_t s);
// Column norms into nrm; every column whose norm is <= rel * max norm refilled with hash-random entries;
// every column scaled to unit norm.  Two launches (norms, then one block per column).
void k_normalize_cols(double* X, int n, int k, double* nrm, double rel, unsigned seed, hipStream_t s);
// The second launch alone, for columns whose norms (256-thread column sums of squares, in col_norms' order)
// are already in nrm.
void k_fill_scale_cols(double* X, int n, int k, const double* nrm, double rel, unsigned seed, hipStream_t s);
// overwrite the columns j with flag[j] != 0 by hash-random entries
void k_random_fill_flagged(double* X, int n, int k, const int* flag, unsigned seed, hipStream_t s);
void k_random_matrix(double* X, size_t n, unsigned seed, hipStream_t s);
void k_init_raw(const double* a, int D, int chi, double* C, double* T, hipStream_t s);
void k_symmetrize_bulk(const double* x, int d, int ph, double* y, hipStream_t s);
void k_symmetrize_bulk_back2(const double* g, const double* x_orig, int d, int ph, double* out,
                             hipStream_t s);
void k_double_layer(const double* bulk, int d, int ph, double* ap, double* a, hipStream_t s);
void k_double_layer_back(const double* bulk, int d, int ph, const double* apbar, const double* abar,
                         double* bbar, hipStream_t s);

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  std::multimap<size_t, void*> free_blocks;   // size -> ptr cache
  std::map<uintptr_t, size_t> live;           // start -> bytes of every block handed out
  size_t cached_bytes = 0;
  // throw unless [p, p + bytes) lies inside one live block
  void check_range(const void* p, size_t bytes, const char* what) const;
  void* alloc(size_t bytes);
  void release(void* p, size_t bytes);
  void trim();
  DevBuf buf(size_t bytes) { return DevBuf(this, bytes); }
  template <typename T> DevBuf bufn(size_t n) { return DevBuf(this, n * sizeof(T)); }
  // Small host <-> device transfers on the stream go through a pinned host staging area (a transfer
  // from or to pageable memory costs tens of microseconds of idle GPU each).  read_async queues a
  // device -> host copy whose data reaches dst_host at the next sync(); write_async copies the host
  // data into the staging area at once and queues the host -> device copy.
  void read_async(void* dst_host, const void* src_dev, size_t bytes);
  void write_async(void* dst_dev, const void* src_host, size_t bytes);
  void sync();   // stream synchronise, then deliver the queued reads
  // device-time phase clocks: record() marks the stream; elapsed(a, b) (ms) is valid once the stream
  // has passed both marks (after a sync)
  hipEvent_t mark(int slot);
  float elapsed(int a, int b);
  ~Ctx();
 private:
  hipEvent_t ev_[8] = {};
  char* pinned_ = nullptr;
  size_t pinned_used_ = 0;
  struct PendingRead { void* dst; size_t off, bytes; };
  std::vector<PendingRead> pending_;
  char* stage(size_t bytes, size_t& off);   // staging space, or nullptr when the transfer is too large
 public:
};

// host <-> device helpers that accept either host or device pointers
void copy_in(Ctx* c, void* dst_dev, const void* src_any, size_t bytes);
void copy_out(Ctx* c, void* dst_any, const void* src_dev, size_t bytes);

}  // namespace tnad
