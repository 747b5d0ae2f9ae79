// Launch-shape table: the built-in shapes of every algorithm and the choice
// among them made by the autotuner (replaces the hard-coded group and grain
// sizes of the reference drivers and bulk::choose_sizes).
#pragma once

#include <stdint.h>

namespace bulkx {

struct LaunchShape {
  int64_t num_groups;
  int group_size;
  int grain_size;
};

// Size classes of the tuning table: n < 2^16, < 2^20, < 2^24, >= 2^24.
constexpr int kSizeClasses = 5;
int size_class(int64_t n);

// Number of tile-shape variants `algo` is compiled with.
int num_variants(int algo);

// Shape of variant v of `algo` over n elements of dtype on the current device.
int variant_shape(int algo, int dtype, int64_t n, int v, LaunchShape* out);

// Variant the tuning table chooses for (algo, dtype, n); 0 (the built-in
// default) when the table has no entry.  The table is read once from
// $BULKX_TUNING or from tuning/mi355x.json next to the library's directory.
int tuned_variant(int algo, int dtype, int64_t n);

}  // namespace bulkx
