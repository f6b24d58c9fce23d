// Size limits of the kernels, checked on the host before anything is launched.  Pure C++ with no HIP
// dependency, so the same checks are compiled into a CPU-only unit test.
//
// Every check returns a Limit: code 0 when the sizes are supported, otherwise the TNAD_ERR_* code the
// entry point returns (TNAD_ERR_SHAPE for sizes that make no sense, TNAD_ERR_UNSUPPORTED for sizes
// beyond what the kernels handle) and the message for tnad_last_error().
#pragma once

namespace tnad {

// the Jacobi eigensolver holds the packed FP64 upper triangle of a 192 x 192 matrix in LDS; the
// subspace eigensolver's block (kept columns plus guard) is limited to the same size
constexpr int JACOBI_MAX = 192;
// largest corner matrix (chi * D) / augmented TRG matrix (d1 d2 + d3 d4) the step buffers are sized for
constexpr long long MAX_N = 8192;

struct Limit {
  int code = 0;             // 0: supported
  const char* msg = "";
  explicit operator bool() const { return code != 0; }
};

constexpr int kErrShape = -2;        // TNAD_ERR_SHAPE
constexpr int kErrUnsupported = -6;  // TNAD_ERR_UNSUPPORTED
constexpr int kErrArg = -1;          // TNAD_ERR_ARG

// CTMRG with environment bond chi and double-layer bond D (corner matrix n = chi * D).  For n above
// JACOBI_MAX the chi kept eigenvectors must fit in the subspace block, which is at most JACOBI_MAX wide.
inline Limit ctmrg_dims_limit(long long D, long long chi) {
  if (D < 1 || chi < 1) return {kErrShape, "bond dimensions must be positive"};
  const long long n = D * chi;
  if (n > MAX_N) return {kErrUnsupported, "chi*D above 8192 is not supported"};
  if (n > JACOBI_MAX && chi > JACOBI_MAX)
    return {kErrUnsupported, "chi above 192 with chi*D above 192 is not supported (kept block > 192)"};
  return {};
}

// ComplexF64 CTMRG: the full spectrum of the complex corner matrix through its real embedding
// (2 chi D <= JACOBI_MAX)
inline Limit zctmrg_dims_limit(long long D, long long chi) {
  if (D < 1 || chi < 1) return {kErrShape, "bond dimensions must be positive"};
  if (2 * D * chi > JACOBI_MAX) return {kErrUnsupported, "complex CTMRG: chi*D above 96 is not supported"};
  return {};
}

// ComplexF64 SVD of an m x p matrix through the embedding of [[0, A], [A^H, 0]] (2 (m + p) <= JACOBI_MAX)
inline Limit zsvd_limit(long long m, long long p) {
  if (m < 1 || p < 1) return {kErrShape, "complex svd: sizes must be positive"};
  if (2 * (m + p) > JACOBI_MAX) return {kErrUnsupported, "complex svd: m + p above 96 is not supported"};
  return {};
}

// ComplexF64 TRG: splits of up to chi^2 x chi^2 through the real embedding of the augmented matrix; the
// eigensolver block (4 chi + 4 columns of the embedding, plus the guard) stays within JACOBI_MAX
inline Limit ztrg_limit(long long q, long long chi, long long niter) {
  if (q < 1 || chi < 1 || niter < 0) return {kErrArg, "complex trg: bad sizes"};
  if (chi > 32) return {kErrUnsupported, "complex trg: chi above 32 is not supported"};
  if (q * q > 1024) return {kErrUnsupported, "complex trg: input bond above 32 is not supported"};
  return {};
}

// environment-only work (the :raw initialisation, the expectation value): no eigensolver involved
inline Limit env_dims_limit(long long D, long long chi) {
  if (D < 1 || chi < 1) return {kErrShape, "bond dimensions must be positive"};
  if (D * chi > MAX_N) return {kErrUnsupported, "chi*D above 8192 is not supported"};
  return {};
}

// tnad_symeig: the want largest-|lambda| pairs of an n x n matrix; above n = JACOBI_MAX the block that
// holds them is at most JACOBI_MAX wide
inline Limit symeig_limit(long long n, long long want) {
  if (n < 1 || want < 1 || want > n) return {kErrArg, "symeig: bad sizes"};
  if (n > MAX_N) return {kErrUnsupported, "symeig: n above 8192 is not supported"};
  if (n > JACOBI_MAX && want > JACOBI_MAX)
    return {kErrUnsupported, "symeig: want above 192 with n above 192 is not supported"};
  return {};
}

// Block width of the top-k eigensolver for n, want and a requested k: the whole spectrum for n <=
// JACOBI_MAX, else k raised to want and capped at min(n, JACOBI_MAX).  Callers check symeig_limit (or
// ctmrg_dims_limit) first, so want never exceeds the result.
inline int eig_block_width(int n, int want, int k) {
  if (n <= JACOBI_MAX) return n;
  int w = k > want ? k : want;
  const int cap = n < JACOBI_MAX ? n : JACOBI_MAX;
  return w < cap ? w : cap;
}

// svd_back for the CTMRG use on an n x n matrix, chi kept columns
inline Limit svd_back_sym_limit(long long n, long long chi) {
  if (n < 1 || chi < 1 || chi > n) return {kErrArg, "svd_back: bad sizes"};
  if (n > MAX_N) return {kErrUnsupported, "svd_back: n above 8192 is not supported"};
  if (n > JACOBI_MAX && chi > JACOBI_MAX)
    return {kErrUnsupported, "svd_back: chi above 192 with n above 192 is not supported"};
  return {};
}

// trg_svd of an m x mp matrix (m = d1 d2, mp = d3 d4) keeping at most dmax triplets: the singular
// triplets come from the augmented (m + mp) x (m + mp) symmetric matrix, whose eigenvalues are +-s, so
// keeping min(dmax, m, mp) triplets needs 2 min(dmax, m, mp) + 2 eigenpairs (the +1 of the keep rule
// and one guard pair); above JACOBI_MAX that block is at most JACOBI_MAX wide.
inline Limit trg_split_limit(long long m, long long mp, long long dmax) {
  if (m < 1 || mp < 1 || dmax < 1) return {kErrArg, "trg_svd: bad sizes"};
  const long long N = m + mp;
  if (N > MAX_N) return {kErrUnsupported, "trg_svd: d1 d2 + d3 d4 above 8192"};
  const long long len = m < mp ? m : mp;
  const long long keep = dmax < len ? dmax : len;
  if (N > JACOBI_MAX && 2 * keep + 2 > JACOBI_MAX)
    return {kErrUnsupported,
            "trg_svd: more than 95 kept triplets with d1 d2 + d3 d4 above 192 is not supported"};
  return {};
}

}  // namespace tnad
