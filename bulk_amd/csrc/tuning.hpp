// Launch-shape table: the built-in shapes of every algorithm and the
// overrides loaded from an autotuner table (replaces the hard-coded group
// and grain sizes of the reference drivers and bulk::choose_sizes).
#pragma once

#include <stdint.h>

namespace bulkx {

struct LaunchShape {
  int64_t num_groups;
  int group_size;
  int grain_size;
};

// Shape of `algo` over n elements of dtype on the current device.
int choose_shape(int algo, int dtype, int64_t n, LaunchShape* out);

// Variant index chosen by the loaded tuning table for (algo, dtype, n), or
// -1 when the table has no entry (the built-in default is used).
int tuned_variant(int algo, int dtype, int64_t n);

}  // namespace bulkx
