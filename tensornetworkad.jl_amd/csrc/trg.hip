// TRG coarse-graining step (src/trg.jl:13-44) and its reverse sweep.
#include "trg.h"

namespace tnad {
void trg_run(Ctx*, const double*, int, int, int, double, double*, double*) {
  throw Error(TNAD_ERR_UNSUPPORTED, "trg: not available in this build");
}
void trg_svd_run(Ctx*, const double*, int, int, int, int, int, double, double*, double*, int*) {
  throw Error(TNAD_ERR_UNSUPPORTED, "trg_svd: not available in this build");
}
}  // namespace tnad
