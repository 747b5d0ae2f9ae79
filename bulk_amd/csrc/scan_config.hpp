// Tile shapes of the single-pass scan: BLOCK threads, ITEMS consecutive
// elements per thread (16 bytes * k per thread so full tiles use dwordx4).
#pragma once

#include <type_traits>

namespace bulkx {

// Variant 0 (the default): 1024-thread workgroups amortise one look-back
// over a 128 KB tile.  Variant 1: a quarter tile (more tiles in flight for
// mid-size inputs).  Variants 2 and 3: the same 128 bytes per thread in 512-
// and 256-thread workgroups, so two or four workgroups share a CU and one
// streams its tile while another scans or looks back.  The tuning table
// picks one per dtype and size class.
// Variant 4: a double tile (256 KB) for 4-byte integers, whose accumulator
// leaves room for 64 elements per thread in a 1024-thread workgroup (the
// float and 8-byte forms would spill, so they run variant 0's shape).
template <class T, int V = 0>
struct ScanConfig {
  static constexpr bool kWide = V == 4 && sizeof(T) == 4 && std::is_integral<T>::value;
  static constexpr int BLOCK = (V == 0 || V == 4) ? 1024 : V == 3 ? 256 : 512;
  static constexpr int ITEMS = kWide ? 64 : (sizeof(T) == 4 ? 32 : 16) / (V == 1 ? 2 : 1);
  static constexpr int TILE = BLOCK * ITEMS;
};
constexpr int kScanVariants = 5;

}  // namespace bulkx
