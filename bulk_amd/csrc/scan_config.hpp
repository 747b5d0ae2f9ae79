// Tile shapes of the single-pass scan: BLOCK threads, ITEMS consecutive
// elements per thread (16 bytes * k per thread so full tiles use dwordx4).
#pragma once

namespace bulkx {

// 1024-thread workgroups amortise one look-back over a 128 KB tile; the
// shapes come from a block x items sweep on MI355X at 2^24 and 2^28 elements.
template <class T>
struct ScanConfig {
  static constexpr int BLOCK = 1024;
  static constexpr int ITEMS = sizeof(T) == 4 ? 32 : 16;
};

}  // namespace bulkx
