// Stable LSD radix sort of (uint32 key, int32 value) pairs on the GPU.
//
// Same result as a stable ascending sort by key (ties keep input order), the
// semantics of thrust::sort_by_key used by the reference
// (src/cuda/force_barnes_hut.cu:324-328). 8-bit digits; per pass:
//   count   - per-tile digit histogram (LDS atomics; the keys of a wave that
//             share its first lane's digit are added by one lane, so skewed
//             digits, as in the top bits of clustered Morton codes, do not
//             serialise on one LDS address)
//   scan    - exclusive scan of each digit's row over the tiles (one workgroup
//             per digit), with the row totals
//   scatter - digit bases from the row totals; in-tile stable ranking by wave
//             ballots (match-any over the digit bits, rank = popcount of lower
//             matching lanes), computed once per key: each key keeps the count
//             of earlier equal digits in its wave's sub-tile, and waves own
//             contiguous sub-tiles so order is kept; the tile is staged in LDS in
//             sorted order, so each digit's run is written to consecutive
//             addresses.
// Tiles of 4,096 keys over 16 waves: at 1M keys, one workgroup per CU, and runs
// of about 16 keys per digit and tile for the output writes (smaller tiles
// write shorter runs and were slower, as were 4- and 8-wave workgroups and
// 10-bit digits). Wave scans are DPP row shifts and row broadcasts.
// The element count is read from device memory, so a sort can follow a kernel
// that produced its length without a host round trip.
#pragma once

#include "common.hpp"

namespace nbx {

struct RadixSort {
  static constexpr int kWaves = 16;
  static constexpr int kRounds = 4;
  static constexpr int kTile = kWaves * kRounds * 64;  // keys per workgroup
  int max_n = 0;
  int max_tiles = 0;
  uint32_t* keys_tmp = nullptr;
  int32_t* vals_tmp = nullptr;
  uint32_t* table = nullptr;  // [256][max_tiles]
  uint32_t* bases = nullptr;  // [256] row totals (keys per digit) of the current pass

  int init(int max_n);
  void release();
  // Sorts keys/vals in place (result back in keys/vals) over bits [0, key_bits).
  // n_dev: device pointer holding the element count (<= max_n).
  int sort(uint32_t* keys, int32_t* vals, const int32_t* n_dev, int key_bits, hipStream_t s);
};

}  // namespace nbx
