// Launch-shape table, the autotuner's choice among the compiled shapes, and
// the choose_sizes replacement (bulk/choose_sizes.hpp:29-75,
// bulk/detail/choose_sizes.inl:32-117).
//
// Each algorithm is compiled with a few tile shapes (variants); a tuning
// table maps (algo, dtype, size class) to the variant measured fastest on the
// device.  The table is JSON written by `python -m bulk_amd.autotune`:
//   {"entries": [{"algo": 6, "dtype": 4, "size_class": 3, "variant": 1, ...}, ...]}
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>

#include "batch_shapes.hpp"
#include "host_common.hpp"
#include "merge_impl.hpp"
#include "reduce_by_key_impl.hpp"
#include "scan_config.hpp"
#include "sort_impl.hpp"
#include "tuning.hpp"

namespace bulkx {

namespace {

constexpr int kAlgos = 11;  // the plain algorithms, then the batch families
constexpr int kDtypes = 6;

struct Table {
  std::mutex mu;
  bool loaded = false;
  // -1 = no entry
  signed char v[kAlgos][kDtypes][kSizeClasses];
  Table() { memset(v, -1, sizeof(v)); }
};

Table& table() {
  static Table t;
  return t;
}

// Integer value of "key": N inside [b, e), or `missing`.
long find_int(const std::string& s, size_t b, size_t e, const char* key, long missing) {
  const std::string pat = std::string("\"") + key + "\"";
  size_t k = s.find(pat, b);
  if (k == std::string::npos || k >= e) return missing;
  size_t c = s.find(':', k + pat.size());
  if (c == std::string::npos || c >= e) return missing;
  return strtol(s.c_str() + c + 1, nullptr, 10);
}

// Parses every {...} object holding the four keys; returns entries stored.
int parse_into(Table& t, const std::string& text) {
  int count = 0;
  size_t pos = 0;
  while (true) {
    size_t k = text.find("\"algo\"", pos);
    if (k == std::string::npos) break;
    size_t b = text.rfind('{', k);
    size_t e = text.find('}', k);
    if (b == std::string::npos || e == std::string::npos) break;
    const long algo = find_int(text, b, e, "algo", -1);
    const long dtype = find_int(text, b, e, "dtype", -1);
    const long sc = find_int(text, b, e, "size_class", -1);
    const long var = find_int(text, b, e, "variant", -1);
    if (algo >= 0 && algo < kAlgos && dtype >= 0 && dtype < kDtypes && sc >= 0 &&
        sc < kSizeClasses && var >= 0 && var < num_variants(int(algo))) {
      t.v[algo][dtype][sc] = static_cast<signed char>(var);
      ++count;
    }
    pos = e + 1;
  }
  return count;
}

int load_file(Table& t, const std::string& path, int* entries) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return BULKX_ERR_INVALID_ARGUMENT;
  std::string text;
  char buf[4096];
  size_t r;
  while ((r = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, r);
  fclose(f);
  memset(t.v, -1, sizeof(t.v));
  const int c = parse_into(t, text);
  if (entries) *entries = c;
  return BULKX_SUCCESS;
}

// tuning/mi355x.json beside the directory that holds libbulkx.so.
std::string default_path() {
  const char* env = getenv("BULKX_TUNING");
  if (env) return std::string(env);
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_path), &info) && info.dli_fname) {
    std::string lib(info.dli_fname);
    size_t slash = lib.rfind('/');
    std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    return dir + "/../tuning/mi355x.json";
  }
  return std::string();
}

void ensure_loaded(Table& t) {
  if (t.loaded) return;
  t.loaded = true;
  const std::string p = default_path();
  if (!p.empty() && p != "none") (void)load_file(t, p, nullptr);
}

}  // namespace

int size_class(int64_t n) {
  if (n < (int64_t(1) << 16)) return 0;
  if (n < (int64_t(1) << 20)) return 1;
  if (n < (int64_t(1) << 24)) return 2;
  if (n < (int64_t(1) << 26)) return 3;
  return 4;
}

int num_variants(int algo) {
  switch (algo) {
    case BULKX_ALGO_INCLUSIVE_SCAN:
    case BULKX_ALGO_EXCLUSIVE_SCAN: return kScanVariants;
    case BULKX_ALGO_REDUCE: return 3;
    case BULKX_ALGO_MERGE:
    case BULKX_ALGO_MERGE_BY_KEY: return kMergeVariants;
    case BULKX_ALGO_SORT_BY_KEY: return kSortVariants;
    case BULKX_ALGO_REDUCE_BY_KEY: return kRbkVariants;
    case BULKX_ALGO_BATCH_SCAN:
    case BULKX_ALGO_BATCH_REDUCE_BY_KEY:
    case BULKX_ALGO_BATCH_MERGE: return kBatchShapes;
    default: return 1;
  }
}

int tuned_variant(int algo, int dtype, int64_t n) {
  if (algo < 0 || algo >= kAlgos || dtype < 0 || dtype >= kDtypes) return 0;
  Table& t = table();
  std::lock_guard<std::mutex> lock(t.mu);
  ensure_loaded(t);
  // the batch families are tuned once per batch, in size class 0
  const int sc = algo >= BULKX_ALGO_BATCH_SCAN ? 0 : size_class(n);
  const int v = t.v[algo][dtype][sc];
  return v < 0 ? 0 : v;
}

int variant_shape(int algo, int dtype, int64_t n, int v, LaunchShape* out) {
  size_t es = dtype_size(dtype);
  if (es == 0) return BULKX_ERR_UNSUPPORTED_DTYPE;
  if (v < 0 || v >= num_variants(algo)) return BULKX_ERR_INVALID_ARGUMENT;
  DeviceProps p;
  int cus = current_device_props(&p) == BULKX_SUCCESS ? p.num_cus : 256;
  LaunchShape s{1, 256, 1};
  switch (algo) {
    case BULKX_ALGO_INCLUSIVE_SCAN:
    case BULKX_ALGO_EXCLUSIVE_SCAN: {
      switch (v) {
#define BULKX_SCAN_SHAPE(V)                                                                 \
  case V:                                                                                   \
    s.group_size = es == 4 ? ScanConfig<float, V>::BLOCK : ScanConfig<double, V>::BLOCK;    \
    s.grain_size = es == 4 ? ScanConfig<float, V>::ITEMS : ScanConfig<double, V>::ITEMS;    \
    break;
        BULKX_SCAN_SHAPE(0)
        BULKX_SCAN_SHAPE(1)
        BULKX_SCAN_SHAPE(2)
        BULKX_SCAN_SHAPE(3)
        case 4:  // 4-byte integers only; the others run variant 0's shape
          s.group_size = 1024;
          s.grain_size = (dtype == BULKX_INT32 || dtype == BULKX_UINT32)
                             ? ScanConfig<int32_t, 4>::ITEMS
                             : es == 4 ? ScanConfig<float, 0>::ITEMS : ScanConfig<double, 0>::ITEMS;
          break;
#undef BULKX_SCAN_SHAPE
      }
      const int64_t tile = int64_t(s.group_size) * s.grain_size;
      s.num_groups = (n + tile - 1) / tile;
      break;
    }
    case BULKX_ALGO_REDUCE: {
      const int per_cu = v == 1 ? 2 : v == 2 ? 6 : 3;
      s.group_size = 256;
      s.grain_size = int(16 / es) * 2;
      const int64_t tile = int64_t(s.group_size) * s.grain_size;
      const int64_t want = (n + tile - 1) / tile;
      s.num_groups = want < int64_t(cus) * per_cu ? want : int64_t(cus) * per_cu;
      break;
    }
    case BULKX_ALGO_REDUCE_BY_KEY: {
      v = rbk_shape(v);
      if (v == 0) {
        s.group_size = RbkConfig<int, int, 0>::BLOCK;
        s.grain_size = RbkConfig<int, int, 0>::ITEMS;
      } else if (v == 1) {
        s.group_size = RbkConfig<int, int, 1>::BLOCK;
        s.grain_size = RbkConfig<int, int, 1>::ITEMS;
      } else if (v == 2) {
        s.group_size = RbkConfig<int, int, 2>::BLOCK;
        s.grain_size = RbkConfig<int, int, 2>::ITEMS;
      } else if (v == 3) {
        s.group_size = RbkConfig<int, int, 3>::BLOCK;
        s.grain_size = RbkConfig<int, int, 3>::ITEMS;
      } else if (v == 4) {
        s.group_size = RbkConfig<int, int, 4>::BLOCK;
        s.grain_size = RbkConfig<int, int, 4>::ITEMS;
      } else {
        s.group_size = RbkConfig<int, int, 5>::BLOCK;
        s.grain_size = RbkConfig<int, int, 5>::ITEMS;
      }
      const int64_t tile = int64_t(s.group_size) * s.grain_size;
      s.num_groups = (n + tile - 1) / tile;
      break;
    }
    case BULKX_ALGO_MERGE:
    case BULKX_ALGO_MERGE_BY_KEY:
    case BULKX_ALGO_SORT_BY_KEY: {
      if (v == 0) {
        s.group_size = es == 4 ? MergeConfig<float, 0>::BLOCK : MergeConfig<double, 0>::BLOCK;
        s.grain_size = es == 4 ? MergeConfig<float, 0>::ITEMS : MergeConfig<double, 0>::ITEMS;
      } else if (v == 1) {
        s.group_size = es == 4 ? MergeConfig<float, 1>::BLOCK : MergeConfig<double, 1>::BLOCK;
        s.grain_size = es == 4 ? MergeConfig<float, 1>::ITEMS : MergeConfig<double, 1>::ITEMS;
      } else if (v == kSortRadixVariant && algo == BULKX_ALGO_SORT_BY_KEY && es == 4) {
        s.group_size = RadixConfig<uint32_t>::BLOCK;  // the radix passes' tile
        s.grain_size = RadixConfig<uint32_t>::ITEMS;
      } else {  // sort variants 2 and 3: the block sort's shape (4-byte keys)
        const bool k4 = es == 4;
        s.group_size = k4 ? (v == 2 ? SortShape<float, float, 2>::BBLOCK
                                    : SortShape<float, float, 3>::BBLOCK)
                          : SortShape<double, double, 0>::BBLOCK;
        s.grain_size = k4 ? SortShape<float, float, 2>::BITEMS : SortShape<double, double, 0>::BITEMS;
      }
      const int64_t tile = int64_t(s.group_size) * s.grain_size;
      s.num_groups = (n + tile - 1) / tile;
      break;
    }
    case BULKX_ALGO_REDUCE_INTERVALS: {
      s.group_size = 256;
      s.grain_size = 1;
      s.num_groups = n;  // one workgroup per interval (capped at 65536 and looped)
      break;
    }
    case BULKX_ALGO_BATCH_SCAN:
    case BULKX_ALGO_BATCH_REDUCE_BY_KEY:
    case BULKX_ALGO_BATCH_MERGE: {
      // the family's tile (one workgroup per tile of a job of n elements)
      const int fam = algo == BULKX_ALGO_BATCH_MERGE ? int(FAM_MERGE)
                      : algo == BULKX_ALGO_BATCH_SCAN ? (dtype == BULKX_FLOAT32 ? int(FAM_SCAN_F32) : int(FAM_SCAN_I32))
                      : dtype == BULKX_FLOAT32 ? int(FAM_RBK_F32)
                      : dtype == BULKX_INT64 ? int(FAM_RBK_I64) : int(FAM_RBK_I32);
      s.group_size = fam_block(fam, v);
      s.grain_size = fam_items(fam, v);
      const int64_t tile = fam_tile(fam, v);
      s.num_groups = (n + tile - 1) / tile;
      break;
    }
    default:
      return BULKX_ERR_INVALID_ARGUMENT;
  }
  if (s.num_groups < 1) s.num_groups = 1;
  *out = s;
  return BULKX_SUCCESS;
}

}  // namespace bulkx

using namespace bulkx;

extern "C" int bulkx_choose_sizes(int algo, int dtype, int64_t n, int64_t* num_groups,
                                  int* group_size, int* grain_size) {
  if (n < 0) return BULKX_ERR_INVALID_ARGUMENT;
  if (algo < 0 || algo >= kAlgos) return BULKX_ERR_INVALID_ARGUMENT;
  LaunchShape s;
  BULKX_TRY(variant_shape(algo, dtype, n, tuned_variant(algo, dtype, n), &s));
  if (num_groups) *num_groups = s.num_groups;
  if (group_size) *group_size = s.group_size;
  if (grain_size) *grain_size = s.grain_size;
  return BULKX_SUCCESS;
}

extern "C" int bulkx_load_tuning(const char* path, int* entries) {
  Table& t = table();
  std::lock_guard<std::mutex> lock(t.mu);
  t.loaded = true;
  const std::string p = path ? std::string(path) : default_path();
  memset(t.v, -1, sizeof(t.v));
  if (entries) *entries = 0;
  if (p.empty() || p == "none") return BULKX_SUCCESS;
  const int st = load_file(t, p, entries);
  // a missing packaged default leaves the built-in shapes; a missing named file is an error
  return (st != BULKX_SUCCESS && path == nullptr) ? BULKX_SUCCESS : st;
}

extern "C" int bulkx_set_tuning(int algo, int dtype, int size_class_, int variant) {
  if (algo < 0 || algo >= kAlgos || dtype < 0 || dtype >= kDtypes || size_class_ < 0 ||
      size_class_ >= kSizeClasses || variant < -1 || variant >= num_variants(algo))
    return BULKX_ERR_INVALID_ARGUMENT;
  Table& t = table();
  std::lock_guard<std::mutex> lock(t.mu);
  ensure_loaded(t);
  t.v[algo][dtype][size_class_] = static_cast<signed char>(variant);
  return BULKX_SUCCESS;
}

extern "C" int bulkx_tuned_variant(int algo, int dtype, int64_t n, int* variant) {
  if (algo < 0 || algo >= kAlgos || dtype < 0 || dtype >= kDtypes || n < 0 || !variant)
    return BULKX_ERR_INVALID_ARGUMENT;
  *variant = tuned_variant(algo, dtype, n);
  return BULKX_SUCCESS;
}

extern "C" int bulkx_tuning_variants(int algo, int* count) {
  if (algo < 0 || algo >= kAlgos || !count) return BULKX_ERR_INVALID_ARGUMENT;
  *count = num_variants(algo);
  return BULKX_SUCCESS;
}
