// ComplexF64 path (zc.hip): planar complex device buffers (real plane, then imaginary plane), complex
// contractions as four real strided contractions, the Hermitian eigensolver through the real embedding,
// and the complex CTMRG energy with its reverse sweep.
#pragma once
#include <map>
#include <string>

#include "tc_build.h"
#include "tnad_internal.h"

namespace tnad {

using ZDims = std::map<char, long long>;   // extent of every index letter of a contraction

// Julia's interleaved ComplexF64 (host or device pointer) <-> planar device buffer of n entries
void z_from_interleaved(Ctx* c, const double* src_any, size_t n, double* dst_planar);
void z_to_interleaved(Ctx* c, const double* src_planar, size_t n, double* dst_any);

// C = alpha op(A) op(B) + beta C for one einsum "ia,ib->ic" (op = conj when the flag is set); planar
void zein(Ctx* c, const std::string& spec, const ZDims& dm, const double* A, bool conjA, const double* B, bool conjB,
          double* Cout, double alpha = 1.0, double beta = 0.0);

// the same on strided views; planeX: entries per plane (offset of the imaginary plane)
void zein_v(Ctx* c, const TView& vA, const double* A, size_t planeA, bool conjA, const TView& vB, const double* B,
            size_t planeB, bool conjB, const TView& vC, double* Cout, size_t planeC, double alpha = 1.0,
            double beta = 0.0);

// Y = X + X^H (n x n planar)
void z_herm_add(Ctx* c, const double* X, int n, double* Y);
// T2 = T1 + conj(T1^{kji})
void z_edge_sym(Ctx* c, const double* T1, int chi, int D, double* T2);
// all n eigenpairs of the Hermitian n x n H (n <= 96), |lambda| descending, phases by the rule of zc.hip:
// U planar n x n, lam (real) n
void z_heig(Ctx* c, const double* H, int n, double* U, double* lam);
// res (K x K, planar) of svd_back (src/trg.jl:76-96): UdU = U^H dU, VdV = V^H dV (K x K, nullable), dS (K, nullable)
void z_svdback_res(Ctx* c, const double* UdU, const double* VdV, const double* dS, const double* sv, int K, double* R);
// X[:, j] *= f[j] (real f, device), planar n x k
void z_scale_cols(Ctx* c, double* X, int n, int k, const double* f);

// energy (src/variationalipeps.jl:28-40) of a complex bulk (planar d x d x d x d x s on the device), h planar
// s x s x s x s; e_host receives e/n (2 doubles); grad (nullable, planar) the gradient of Re(energy).
void zenergy_run(Ctx* c, const double* h, const double* bulk, int d, int ph, int chi, double tol, int maxit,
                 double* e_host, double* grad, int& iters);

// thin SVD of a complex m x p matrix (m + p <= 96), planar: U m x K, S K (real), V p x K, K = min(m, p)
void zsvd_small(Ctx* c, const double* A, int m, int p, double* U, double* S, double* V);
// its svd_back (src/trg.jl:72-106): dU, dS, dV nullable; dA m x p
void zsvd_back_small(Ctx* c, const double* A, int m, int p, const double* dU, const double* dS, const double* dV,
                     double* dA);

// trg (src/trg.jl:13-30) of a complex q^4 tensor (planar, device): lnz_out (2 doubles) = lnZ per site; abar_out
// (nullable, planar q^4) the gradient of Re(lnZ)
void ztrg_run(Ctx* c, const double* a_planar, int q, int chi, int niter, double tol, double* lnz_out,
              double* abar_out);

}  // namespace tnad
