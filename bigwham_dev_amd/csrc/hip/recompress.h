// Recompression of the low-rank factors (see recompress.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/bigwham_flat.h"

namespace bwl {

// A flat H-matrix that borrows h's full-rank blocks and permutations and
// owns the recompressed low-rank factors (blocks whose rank did not drop
// point at h's factors, so h must outlive it).
struct RecompressedFlat {
  bw_flat_hmat flat{};
  std::vector<int64_t> lr_pattern;          // row 5 holds the new ranks
  std::vector<std::vector<double>> a, b;    // new factors (empty if unchanged)
  std::vector<const double*> pa, pb;
  int64_t entries_before = 0, entries_after = 0;  // low-rank entries
  int64_t blocks_changed = 0;
};

// Every low-rank block A B of h replaced by the shortest A' B' with
// ||A B - A' B'||_F <= tol ||A B||_F (0 < tol < 1).
// expected_area: the point pairs the blocks must cover (-1: the whole
// matrix; a block-diagonal composite covers only its diagonal blocks).
int recompress_flat(const bw_flat_hmat* h, double tol, RecompressedFlat* out, std::string* msg,
                    long double expected_area = -1.0L);

}  // namespace bwl
