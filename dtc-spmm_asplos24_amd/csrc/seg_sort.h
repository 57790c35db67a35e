// Segmented device-wide pieces used by the ME-TCF build (implemented in
// primitives.hip).  Kept apart from dtc_internal.h, which the SpMM kernels
// include, so that these declarations do not touch the SpMM sources.
#pragma once

#include "dtc_internal.h"

namespace dtc {

// Segmented variant of radix_sort_pairs over 32-bit keys: the array is
// segments [csb[w'], csb[w'+1]), each cut into its own tiles of
// seg_radix_tile_size() keys (tile t of segment tw[t]; segment w' owns tiles
// [tb[w'], tb[w'+1]), *ntiles_dev in all, at most ntiles); every segment is
// sorted stably by key bits [0, key_bits) on its own, in place of the whole
// array.
size_t seg_radix_sort_workspace_bytes(int64_t n, int64_t segments);
int64_t seg_radix_tile_size();
dtc_status_t seg_radix_sort_pairs_u32(uint32_t *keys0, int32_t *vals0, uint32_t *keys1,
                                      int32_t *vals1, int key_bits, const int32_t *tw,
                                      const int32_t *tb, const int32_t *csb,
                                      const int32_t *ntiles_dev, int64_t ntiles,
                                      bool *result_in_1, void *ws, hipStream_t stream);
// scan_key_changes over 32-bit keys.
dtc_status_t scan_key_changes_u32(const uint32_t *keys, int32_t *out, int64_t n, void *ws,
                                  hipStream_t stream);

}  // namespace dtc
