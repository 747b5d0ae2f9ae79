// Tile shapes of the single-pass scan: BLOCK threads, ITEMS consecutive
// elements per thread (16 bytes * k per thread so full tiles use dwordx4).
#pragma once

namespace bulkx {

template <class T>
struct ScanConfig {
  static constexpr int BLOCK = 256;
  static constexpr int ITEMS = sizeof(T) == 4 ? 32 : 16;
};

}  // namespace bulkx
