// TRG (src/trg.jl:13-44) on the device: host orchestration entry points.
#pragma once
#include "tnad_internal.h"

namespace tnad {
void trg_run(Ctx* c, const double* a, int q, int chi, int niter, double tol, double* lnz, double* abar);
void trg_svd_run(Ctx* c, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol, double* u,
                 double* v, int* k_out);
void trg_svd_back_run(Ctx* c, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol,
                      const double* ubar, const double* vbar, int k, double* tbar);
}  // namespace tnad
