// Host-only check of the size limits (csrc/tnad_limits.h) that every entry point applies before it launches a
// kernel: the largest supported shapes pass, the first unsupported ones fail with the right code, and the
// eigensolver block width always holds the requested columns.
#include <cstdio>

#include "tnad_limits.h"

using namespace tnad;

static int bad = 0;
#define EXPECT(cond)                                                   \
  do {                                                                 \
    if (!(cond)) {                                                     \
      std::printf("FAIL line %d: %s\n", __LINE__, #cond);              \
      ++bad;                                                           \
    }                                                                  \
  } while (0)

int main() {
  // CTMRG: chi * D <= 192 takes the full spectrum, any chi
  EXPECT(!ctmrg_dims_limit(1, 192));
  EXPECT(!ctmrg_dims_limit(16, 12));
  // n > 192: the chi kept columns must fit the 192-wide block
  EXPECT(!ctmrg_dims_limit(16, 128));                            // chi = 128, n = 2048 (the bench shape)
  EXPECT(!ctmrg_dims_limit(2, 192));                             // n = 384, chi = 192
  EXPECT(ctmrg_dims_limit(2, 193).code == kErrUnsupported);      // n = 386, chi = 193
  EXPECT(ctmrg_dims_limit(1, 193).code == kErrUnsupported);      // n = 193 = chi: block too narrow
  EXPECT(ctmrg_dims_limit(16, 513).code == kErrUnsupported);     // n = 8208 > 8192
  EXPECT(ctmrg_dims_limit(0, 4).code == kErrShape);
  EXPECT(ctmrg_dims_limit(4, 0).code == kErrShape);
  EXPECT(!env_dims_limit(4, 1000));
  EXPECT(env_dims_limit(16, 1000).code == kErrUnsupported);
  // symeig
  EXPECT(!symeig_limit(192, 192));
  EXPECT(!symeig_limit(2048, 192));
  EXPECT(symeig_limit(2048, 193).code == kErrUnsupported);
  EXPECT(symeig_limit(200, 201).code == kErrArg);
  EXPECT(symeig_limit(8193, 10).code == kErrUnsupported);
  EXPECT(!svd_back_sym_limit(2048, 128));
  EXPECT(svd_back_sym_limit(2048, 193).code == kErrUnsupported);
  // block widths always hold the wanted columns whenever the limit passes
  int widths = 0;
  for (int n = 1; n <= 600; ++n)
    for (int want = 1; want <= n; want += 7)
      for (int k = 0; k <= 220; k += 11) {
        if (symeig_limit(n, want)) continue;
        const int w = eig_block_width(n, want, k);
        EXPECT(w >= want && w <= (n < JACOBI_MAX ? n : JACOBI_MAX));
        if (n <= JACOBI_MAX) EXPECT(w == n);
        ++widths;
      }
  // TRG split: 2 min(dmax, m, mp) + 2 eigenpairs of the augmented matrix
  EXPECT(!trg_split_limit(400, 400, 64));    // chi = 20 TRG: (20 x 20) x (20 x 20) splits
  EXPECT(!trg_split_limit(4096, 4096, 64));  // chi = 64
  EXPECT(!trg_split_limit(144, 144, 95));
  EXPECT(trg_split_limit(144, 144, 96).code == kErrUnsupported);   // (12,12,12,12), dmax 96
  EXPECT(trg_split_limit(144, 144, 64).code == 0);
  EXPECT(!trg_split_limit(90, 90, 200));     // N = 180 <= 192: full spectrum, any dmax
  EXPECT(trg_split_limit(4096, 4097, 1).code == kErrUnsupported);
  EXPECT(trg_split_limit(0, 4, 1).code == kErrArg);
  std::printf("%d block widths checked, %d bad\n", widths, bad);
  return bad ? 1 : 0;
}
