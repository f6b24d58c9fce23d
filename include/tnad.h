/*
 * tnad.h -- C ABI of libtnad: CTMRG and TRG renormalisation steps of
 * TensorNetworkAD.jl and their reverse-mode adjoint, as hand-written HIP
 * kernels for AMD Instinct MI355X (gfx950), one GPU per context.
 *
 * Conventions
 *  - Arrays are column-major (Julia memory order), FP64, passed as plain
 *    pointers.  Every array argument may be a host pointer or a device
 *    pointer on the context's GPU; the library detects which.
 *  - Index names and orders follow the reference: the bulk (double layer)
 *    tensor a[j,k,l,m] is D x D x D x D; the corner C is chi x chi; the edge
 *    T[i,j,k] is chi x D x chi; the iPEPS bulk is d x d x d x d x s.
 *  - Calls are synchronous on return.  A context owns one HIP stream and is
 *    not thread safe; distinct contexts are independent.
 *  - Return value: TNAD_OK (0) or a negative TNAD_ERR_* code; the message is
 *    available from tnad_last_error().  Non-convergence is not an error: it is
 *    reported through *iters == maxit, as the reference's fixedpoint returns
 *    silently at maxit (src/fixedpoint.jl:33).
 *
 * Size limits (checked on the host before any kernel runs; TNAD_ERR_UNSUPPORTED otherwise):
 *  - CTMRG / energy: chi*D <= 8192, and chi <= 192 whenever chi*D > 192 (the chi kept eigenvectors
 *    must fit the eigensolver's block of at most 192 columns).
 *  - tnad_symeig: n <= 8192, and want <= 192 whenever n > 192.
 *  - tnad_trg_svd / tnad_trg_svd_back: d1 d2 + d3 d4 <= 8192, and min(dmax, d1 d2, d3 d4) <= 95 whenever
 *    d1 d2 + d3 d4 > 192; tnad_trg: chi <= 64.
 *
 * The reference (TensorNetworkAD.jl) has no FFI; the functions below replace
 * the Julia methods cited next to each declaration (paths relative to the
 * reference repository).  The Julia-side ccall bindings and ChainRulesCore
 * rrules that attach them are listed in INTEGRATION.md.
 */
#ifndef TNAD_H
#define TNAD_H

#ifdef __cplusplus
extern "C" {
#endif

#define TNAD_ABI_VERSION 1

typedef struct tnad_ctx tnad_ctx;
typedef struct tnad_tape tnad_tape;

enum {
  TNAD_OK = 0,
  TNAD_ERR_ARG = -1,         /* invalid argument / null pointer */
  TNAD_ERR_SHAPE = -2,       /* DimensionMismatch in the reference (src/ipeps.jl:19-20) */
  TNAD_ERR_HIP = -3,         /* HIP runtime error */
  TNAD_ERR_NAN = -4,         /* NaN detected in singular values / residuals */
  TNAD_ERR_ALLOC = -5,       /* device allocation failed */
  TNAD_ERR_UNSUPPORTED = -6  /* size outside what the kernels support */
};

enum { TNAD_PREC_FP64 = 0, TNAD_PREC_FP32 = 1, TNAD_PREC_BF16 = 2 };

/* Run options of the CTMRG fixed point.  tnad_opts_default() gives the
 * reference semantics: FP64 contractions, cold-started eigensolver every step,
 * full unrolled reverse sweep. */
typedef struct tnad_opts {
  int precision;       /* contraction precision of the low-precision phase (TNAD_PREC_*) */
  double switch_tol;   /* precision != FP64: switch to FP64 once |vals - oldvals| <= switch_tol */
  int bwd_precision;   /* contraction precision of the reverse sweep */
  int eig_block;       /* subspace block size k for chi*D > 192 (0 = 2.25 chi, capped at 192) */
  int eig_max_iters;   /* cap on subspace iterations per step, cold start (stops at eig_tol) */
  int eig_warm;        /* 1: start each step's eigensolver from the previous step's Ritz vectors */
  int eig_warm_iters;  /* cap on subspace iterations per step, warm start (stops at eig_tol) */
  double eig_tol;      /* relative Ritz residual required of the kept chi vectors in FP64 steps;
                          low-precision steps use max(eig_tol, 1e-3 switch_tol) */
  int bwd_last;        /* reverse sweep over the last bwd_last steps only (0 = all steps) */
} tnad_opts;

/* Per-call statistics of the last tnad_ctmrg / tnad_energy call. */
typedef struct tnad_stats {
  int iters;            /* CTMRG steps taken */
  int low_prec_iters;   /* of which at the low contraction precision */
  int eig_iters_total;  /* subspace iterations summed over steps */
  int eig_k;            /* eigensolver block size (n when the full spectrum is computed) */
  double eig_resid_max; /* largest relative Ritz residual over the steps */
  int bwd_steps;        /* steps covered by the last reverse sweep */
  double fwd_ms;        /* device time of the CTMRG steps (HIP events on the context's stream) */
  double eig_ms;        /* of which in the steps' eigensolver */
  double bwd_ms;        /* device time of the reverse sweep (expectation value and steps), 0 without one */
} tnad_stats;

int tnad_version(void);
void tnad_opts_default(tnad_opts* opts);

/* context: one GPU, one stream */
int tnad_create(int device, tnad_ctx** ctx);
int tnad_destroy(tnad_ctx* ctx);
const char* tnad_last_error(const tnad_ctx* ctx);
int tnad_get_stats(const tnad_ctx* ctx, tnad_stats* out);

/* indexperm_symmetrize (src/ipeps.jl:32-39): bulk d x d x d x d x s -> out (same shape) */
int tnad_indexperm_symmetrize(tnad_ctx* ctx, const double* bulk, int d, int s, double* out);

/* double layer (src/variationalipeps.jl:30-34): ap (D,D,D,D,s,s) and a = tr_phys ap (D,D,D,D), D = d^2.
 * ap may be NULL. */
int tnad_double_layer(tnad_ctx* ctx, const double* bulk, int d, int s, double* ap, double* a);

/* _initializect_square(bulk, Val(:raw), chi) (src/ctmrg.jl:74-86) */
int tnad_init_raw(tnad_ctx* ctx, const double* a, int D, int chi, double* C, double* T);

/* svd(cpmat) on the symmetric n x n matrix A (src/ctmrg.jl:136): the k eigenpairs of largest |lambda|,
 * sorted by |lambda| descending (n <= 192: all n, exactly).  evals: k, evecs: n x k. Returns k via *k_out. */
int tnad_symeig(tnad_ctx* ctx, const double* A, int n, int want, int k, double* evals, double* evecs,
                int* k_out);

/* tnad_symeig for n <= 192 (all n eigenpairs, sorted by |lambda| descending) with no allocation: no
 * counterpart in the reference; for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  A (n x n), evals (n) and evecs (n x n) must be device memory on the context's GPU; the
 * work space (at most 1.20 MB, n = 192) is a file-scope device array of which each GPU holds one copy
 * from load to process exit.  Calls that overlap on one GPU (two contexts used from two threads) race on
 * that array.  n > 192: TNAD_ERR_UNSUPPORTED. */
int tnad_symeig_noalloc(tnad_ctx* ctx, const double* A, int n, double* evals, double* evecs);

/* svd_back (src/trg.jl:72-106) for the CTMRG use: A symmetric n x n, cotangent G (n x chi) of the first chi
 * left singular vectors, dS = dV = nothing; returns Abar + Abar^T (the cotangent of cpmat before
 * cpmat += cpmat', src/ctmrg.jl:135). */
int tnad_svd_back_sym(tnad_ctx* ctx, const double* A, int n, int chi, int k, const double* G, double* Abar);

/* ctmrgstep (src/ctmrg.jl:126-153): one step; vals has chi*D entries, s ./ s[1] sorted descending:
 * all of them when chi*D <= 192 (full spectrum, as the reference), otherwise the k largest followed by
 * zeros, k = eig_block or 2.25 chi capped at 192 (the chi kept values, resolved to the eigensolver
 * tolerance, then the eigensolver's guard block, whose Ritz values are approximate). */
int tnad_ctmrgstep(tnad_ctx* ctx, const double* a, int D, int chi, const double* C, const double* T,
                   const tnad_opts* opts, double* Cout, double* Tout, double* vals);

/* ctmrg (src/ctmrg.jl:110-117 with fixedpoint/StopFunction, src/fixedpoint.jl:11-41).
 * C0/T0: initial environment (e.g. from tnad_init_raw).  The StopFunction counter starts at -1
 * (src/ctmrg.jl:114), so at most maxit + 1 steps run (maxit = 0: one step); a run stops earlier once
 * |vals - oldvals|_2 <= tol.  vals is as for tnad_ctmrgstep.  For chi*D > 192 the stop norm is a lower
 * bound of the reference's: the change of the chi kept values, plus, for the values beyond them, the
 * change of their norm |cpmat - Z Theta Z^T|_F / s[1] less its error bound.  A run therefore never
 * stops after the reference's, and stops before it only when the discarded values move by more than
 * their norm shows (over 26 oracle runs at D = 3, chi = 24 and 64: 24 equal, 2 earlier).
 * If tape != NULL a checkpoint of every step is kept on the device for tnad_ctmrg_back; free it with
 * tnad_tape_free. */
int tnad_ctmrg(tnad_ctx* ctx, const double* a, int D, int chi, const double* C0, const double* T0, double tol,
               int maxit, const tnad_opts* opts, double* C, double* T, double* vals, int* iters,
               tnad_tape** tape);

/* reverse sweep of tnad_ctmrg (what Zygote computes through the unrolled fixedpoint loop):
 * cotangents of the final C, T -> cotangent of a.  The initial environment gets no gradient
 * (@nograd _initializect_square, src/autodiff.jl:5). */
int tnad_ctmrg_back(tnad_ctx* ctx, tnad_tape* tape, const double* Cbar, const double* Tbar, double* abar);
int tnad_tape_free(tnad_tape* tape);

/* expectationvalue (src/variationalipeps.jl:49-56) */
int tnad_expectationvalue(tnad_ctx* ctx, const double* h, int s, const double* ap, int D, int chi,
                          const double* C, const double* T, double* e);

/* energy (src/variationalipeps.jl:28-40) and, if grad != NULL, its gradient with respect to the
 * (unsymmetrised) bulk -- one boundary crossing per objective + gradient evaluation. */
int tnad_energy(tnad_ctx* ctx, const double* h, const double* bulk, int d, int s, int chi, double tol,
                int maxit, const tnad_opts* opts, double* e, double* grad, int* iters);

/* ---------------------------------------------------------------------------------------------
 * ComplexF64.  Complex arrays are Julia's ComplexF64 memory: interleaved (re, im) pairs of doubles,
 * column-major; `double*` points at the first real part.  Every complex contraction runs as four
 * FP64 matrix-core contractions on split real / imaginary planes; Hermitian eigenproblems go through
 * the real symmetric embedding [[Re H, -Im H], [Im H, Re H]] and the FP64 Jacobi solver, so the
 * complex CTMRG takes chi*D <= 96 and the complex SVD m + p <= 96 (TNAD_ERR_UNSUPPORTED beyond).
 * Singular / eigenvector phases: column x is multiplied by conj(q)/|q|, q = sum_r w_r x_r with
 * w_r = 1 + frac(0.6180339887 r) (the first entry of largest modulus made real positive when
 * |q| <= 1e-6 |x|); real data stays real.  The reference leaves these phases to LAPACK, and its
 * energy of a complex bulk depends on them (src/variationalipeps.jl:53-54 contracts l with l).
 * ------------------------------------------------------------------------------------------- */

/* energy (src/variationalipeps.jl:28-40) of a complex bulk (d x d x d x d x s ComplexF64) with a complex
 * h (s x s x s x s): e receives e/n (2 doubles, re and im) and, if grad != NULL, the gradient of Re(energy)
 * (what optimiseipeps differentiates, src/variationalipeps.jl:71) in Zygote's convention
 * dRe(e)/dRe(x) + i dRe(e)/dIm(x) (d x d x d x d x s ComplexF64).  FP64 throughout; the reference's
 * StopFunction semantics with the full spectrum. */
int tnad_energy_z(tnad_ctx* ctx, const double* h, const double* bulk, int d, int s, int chi, double tol, int maxit,
                  double* e, double* grad, int* iters);

/* trg (src/trg.jl:13-30) of a complex q x q x q x q tensor: lnz receives ln Z per site (2 doubles, re and im:
 * the log of the final trace is complex) and, if abar != NULL, the gradient of Re(lnZ) with respect to a
 * (q^4 ComplexF64, Zygote's convention), as test/trg.jl:22-28 differentiates real(trg(...)).  chi <= 32. */
int tnad_trg_z(tnad_ctx* ctx, const double* a, int q, int chi, int niter, double tol, double* lnz, double* abar);

/* svd (LinearAlgebra.svd, thin) of a complex m x p matrix A: U (m x K), S (K, real, descending),
 * V (p x K) with A = U diag(S) V^H, K = min(m, p). */
int tnad_svd_z(tnad_ctx* ctx, const double* A, int m, int p, double* U, double* S, double* V);

/* svd_back (src/trg.jl:72-106) for the SVD tnad_svd_z computes for the same A: cotangents dU (m x K),
 * dS (K, real) and dV (p x K), each nullable (all NULL: a zero dA, as the reference returns nothing,
 * :73), including the i Im diag(U^H dU) Sinv term (:85) and the projector terms of a non-square A
 * (:98-104); writes dA (m x p). */
int tnad_svd_back_z(tnad_ctx* ctx, const double* A, int m, int p, const double* dU, const double* dS,
                    const double* dV, double* dA);

/* trg (src/trg.jl:13-30): ln Z per site of the rank-4 tensor a (q x q x q x q); if abar != NULL also
 * d lnZ / d a (q^4 entries). */
int tnad_trg(tnad_ctx* ctx, const double* a, int q, int chi, int niter, double tol, double* lnz,
             double* abar);

/* trg_svd (src/trg.jl:33-44) on a (d1,d2,d3,d4) tensor: u (d1,d2,k), v (k,d3,d4); *k_out = kept count.
 * u and v must hold min(dmax, d1*d2, d3*d4) columns/rows. */
int tnad_trg_svd(tnad_ctx* ctx, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol,
                 double* u, double* v, int* k_out);

/* Pullback of trg_svd: the svd adjoint of src/trg.jl:55-61,72-106 (with the non-square projector terms of
 * :98-104 for d1*d2 != d3*d4) applied to the cotangents ubar (d1,d2,k) and vbar (k,d3,d4); writes tbar
 * (d1,d2,d3,d4).  k must be the kept count tnad_trg_svd returned for the same t, dmax and tol. */
int tnad_trg_svd_back(tnad_ctx* ctx, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol,
                      const double* ubar, const double* vbar, int k, double* tbar);

/* One strided contraction C = A * B given as an einsum over single-character index names
 * (e.g. "ibcl", "jkcb" -> "ijlk"; the contraction kernel behind every ein"..." call site of the
 * step).  Tensors are contiguous column-major FP64 with the listed extents; precision selects the
 * matrix-core path (FP64, FP32, or BF16 operands with FP32 accumulation). */
int tnad_contract(tnad_ctx* ctx, int precision, const char* ia, const long long* da, const double* A,
                  const char* ib, const long long* db, const double* B, const char* ic, const long long* dc,
                  double alpha, double beta, double* Cout);   /* Cout = alpha A*B + beta Cout */

/* Average device time of one launch of the contraction described as in tnad_contract, measured with
 * HIP events recorded on the context's stream around `reps` back-to-back launches on device-resident
 * operands (no transfers inside the timed region).  For roofline measurement. */
int tnad_time_contract(tnad_ctx* ctx, int precision, const char* ia, const long long* da, const char* ib,
                       const long long* db, const char* ic, const long long* dc, int reps, double* avg_ms);

/* Average device time of one launch of the Jacobi eigensolver's rotation kernel (one workgroup, the packed
 * m x m FP64 matrix in LDS; the warm-start codec's solver) on a random symmetric m x m
 * matrix, `sweeps` sweeps unless it converges sooner (*rounds, nullable, receives the rounds one launch
 * ran), HIP events on the context's stream over `reps` launches.  2 <= m <= 192, 1 <= sweeps <= 60.  For
 * roofline measurement; no counterpart in the reference. */
int tnad_time_jacobi(tnad_ctx* ctx, int m, int sweeps, int reps, double* avg_ms, int* rounds);

/* Roofline timing of the divide-and-conquer symmetric eigensolver (no counterpart in the reference): its
 * stages on a random symmetric m x m matrix (3 <= m <= 192), averaged over reps runs after one warm-up, HIP
 * events on the context's stream: stage_ms[0] Householder tridiagonalisation (one workgroup), stage_ms[1]
 * divide and conquer on the tridiagonal matrix (one workgroup, then the top merge's roots and z over the
 * chip), stage_ms[2] back-transformation, the top merge's product and the sort. */
int tnad_time_syev(tnad_ctx* ctx, int m, int reps, double* stage_ms);

/* ---------------------------------------------------------------------------------------------
 * Learned warm start.  The reference starts CTMRG from _initializect_square (src/ctmrg.jl:62-86,
 * :raw or :random, @nograd per src/autodiff.jl:5) and has no learned counterpart; these entries
 * let a caller start tnad_ctmrg / a CTMRG run from a predicted environment instead.  CTMRG still
 * runs to the caller's tolerance, so the fixed point and its gradient are unchanged.
 * ------------------------------------------------------------------------------------------- */
typedef struct tnad_mlp tnad_mlp;

/* Fully connected network n_in -> hidden (tanh) x n_layers -> n_out (linear), FP32 parameters on
 * the context's GPU, weights uniform(+-1/sqrt(fan_in)) from `seed`, biases 0.  Parameter order:
 * W1 (hidden x n_in, column-major), b1, ..., W_out (n_out x hidden), b_out.  Free with
 * tnad_mlp_free before destroying the context. */
int tnad_mlp_create(tnad_ctx* ctx, int n_in, int hidden, int n_layers, int n_out, unsigned long long seed,
                    tnad_mlp** mlp);
int tnad_mlp_free(tnad_mlp* mlp);
int tnad_mlp_num_params(const tnad_mlp* mlp, long long* n);
int tnad_mlp_get_params(tnad_ctx* ctx, const tnad_mlp* mlp, float* theta);
int tnad_mlp_set_params(tnad_ctx* ctx, tnad_mlp* mlp, const float* theta);   /* also resets Adam */
/* Y (n_out x B) = network(X (n_in x B)) */
int tnad_mlp_forward(tnad_ctx* ctx, const tnad_mlp* mlp, const float* X, int B, float* Y);
/* loss = mean_b 0.5 |network(x_b) - t_b|^2 and (grad != NULL) its gradient (num_params floats) */
int tnad_mlp_loss_grad(tnad_ctx* ctx, tnad_mlp* mlp, const float* X, const float* T, int B, double* loss,
                       float* grad);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; *loss (nullable) is the loss before the step */
int tnad_mlp_train_step(tnad_ctx* ctx, tnad_mlp* mlp, const float* X, const float* T, int B, float lr,
                        double* loss);

/* Network input of a bulk (d x d x d x d x s): the double layer a of the symmetrised bulk
 * (src/ipeps.jl:32-39, src/variationalipeps.jl:30-34), normalised; D^4 floats with D = d^2. */
int tnad_bulk_features(tnad_ctx* ctx, const double* bulk, int d, int s, float* feat);
/* Canonical-gauge features of an environment (C chi x chi, T chi x D x chi, chi <= 192):
 * C = O diag(lam) O^T (|lam| descending), T~ = O^T T O with each column of O signed so that
 * sum_{r<i,a} (a+1)/(r+1) T~[r,a,i] >= 0 (column 0: largest-|.| component positive), then
 * lam[0] >= 0 and sum_a (a+1) T~[0,a,0] >= 0 (global signs of C and T);
 * feat = [lam ; T~] (chi + chi*D*chi floats).  Invariant under C -> S^T C S, T -> S^T T S. */
int tnad_env_features(tnad_ctx* ctx, const double* C, const double* T, int chi, int D, float* feat);
/* Starting environment from features: C0 = diag(lam), T0 = (T~ + T~^{kji}) / 2, both normalised. */
int tnad_env_from_features(tnad_ctx* ctx, const float* feat, int chi, int D, double* C0, double* T0);

#ifdef __cplusplus
}
#endif
#endif /* TNAD_H */
