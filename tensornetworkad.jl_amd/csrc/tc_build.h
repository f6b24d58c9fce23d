// Einsum-string front end for the strided contraction kernel.
//
// A tensor view names each index with one character and carries extents and
// element strides (column-major by default).  contract_desc() classifies the
// indices of  C = A * B  into batch (in A, B, C), M (A, C), N (B, C) and
// K (A, B), merges index runs that are contiguous in every tensor, and emits
// the TcDesc for one launch.  Call sites therefore read like the reference's
// ein"..." strings.  Pure host C++; no HIP dependency.
#pragma once
#include <string>
#include <vector>

#ifndef TNAD_TC_DESC_DEFINED
#include "tnad_internal.h"
#endif

namespace tnad {

struct TView {
  std::string idx;
  std::vector<long long> dim;
  std::vector<long long> str;
};

// column-major contiguous view
TView cm(const std::string& idx, const std::vector<long long>& dims);
// same memory, indices renamed (e.g. a matrix view of a 4-tensor is built with cm())
TView with_strides(const std::string& idx, const std::vector<long long>& dims,
                   const std::vector<long long>& strides);

TcDesc contract_desc(const TView& A, const TView& B, const TView& C, double alpha = 1.0,
                     double beta = 0.0);

}  // namespace tnad
