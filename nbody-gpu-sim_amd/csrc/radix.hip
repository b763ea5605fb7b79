// Stable LSD radix sort (see radix.hpp).
#include "radix.hpp"

namespace nbx {

namespace {

constexpr int kBins = 256;
constexpr int kThreads = RadixSort::kWaves * 64;
static_assert(kThreads >= kBins, "one thread per digit in the digit-wise steps");

__global__ __launch_bounds__(kThreads) void radix_count(const uint32_t* __restrict__ keys,
                                                        const int32_t* __restrict__ n_dev,
                                                        int shift, uint32_t* table, int tiles) {
  constexpr int PER = RadixSort::kTile / kThreads;
  __shared__ uint32_t hist[kBins];
  const int n = *n_dev;
  if (threadIdx.x < kBins) hist[threadIdx.x] = 0;
  const int base = blockIdx.x * RadixSort::kTile;
  uint32_t k[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = base + r * kThreads + threadIdx.x;
    k[r] = i < n ? keys[i] : 0u;
  }
  __syncthreads();
  // the wave's keys that share the first valid lane's digit are counted by one
  // lane (skewed digit distributions would serialise on one LDS address)
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int i = base + r * kThreads + threadIdx.x;
    const bool valid = i < n;
    const uint32_t dg = (k[r] >> shift) & (kBins - 1);
    const uint64_t vm = __ballot(valid);
    if (vm) {
      const int fl = (int)__builtin_ctzll(vm);
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)dg, fl);
      const uint64_t same = __ballot(valid && dg == d0);
      if ((threadIdx.x & 63) == fl) atomicAdd(&hist[d0], (uint32_t)__popcll(same));
      else if (valid && dg != d0) atomicAdd(&hist[dg], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < kBins) table[threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Inclusive scan over the wave by DPP: row shifts 1, 2, 4, 8 within each row
// of 16 lanes, then row broadcasts of lanes 15 and 31.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
  return x;
}

// Exclusive scan of each digit's row of the digit-major table over the tiles,
// one workgroup per digit, recording the row totals; every scatter workgroup
// turns the 256 totals into digit bases itself.
__global__ __launch_bounds__(256) void radix_scan_rows(uint32_t* table, int tiles,
                                                       uint32_t* row_total) {
  __shared__ uint32_t wsum[4];
  uint32_t* row = table + (size_t)blockIdx.x * tiles;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (int base = 0; base < tiles; base += 256) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < tiles ? row[i] : 0u;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = carry, tot = 0;
    for (int q = 0; q < 4; ++q) {
      if (q < w) before += wsum[q];
      tot += wsum[q];
    }
    if (i < tiles) row[i] = before + incl - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) row_total[blockIdx.x] = carry;
}

// Exclusive scan of one value per thread over the workgroup (kThreads).
__device__ int block_scan_u32(uint32_t v, int* total) {
  __shared__ uint32_t ws[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) ws[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (int q = 0; q < kThreads / 64; ++q) {
    if (q < w) before += ws[q];
    tot += ws[q];
  }
  *total = (int)tot;
  return (int)(before + incl - v);
}

__device__ __forceinline__ uint64_t match_digit(uint32_t digit, bool valid, int nbits) {
  uint64_t m = __ballot(valid);
  for (int bit = 0; bit < nbits; ++bit) {
    const bool b = (digit >> bit) & 1u;
    const uint64_t v = __ballot(b);
    m &= b ? v : ~v;
  }
  return m;
}

__global__ __launch_bounds__(kThreads) void radix_scatter(
    const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    int32_t* __restrict__ vout, const uint32_t* __restrict__ table,
    const uint32_t* __restrict__ row_total, const int32_t* __restrict__ n_dev, int shift,
    int nbits, int tiles) {
  constexpr int R = RadixSort::kRounds;
  constexpr int W = RadixSort::kWaves;
  constexpr int T = RadixSort::kTile;
  __shared__ uint32_t wave_hist[W][kBins];
  __shared__ uint32_t local_start[kBins];  // first slot of each digit in the tile's sorted order
  __shared__ uint32_t gstart[kBins];       // first output position of each digit for this tile
  __shared__ uint32_t skey[T];
  __shared__ int32_t sval[T];
  const int n = *n_dev;
  const int base = blockIdx.x * T;
  if (base >= n) return;
  const int tile_n = min(T, n - base);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  for (int k = threadIdx.x; k < W * kBins; k += kThreads) (&wave_hist[0][0])[k] = 0;
  // this tile's digit offsets and the digit totals, loaded before the ranking
  const bool isd = threadIdx.x < kBins;
  const uint32_t my_row_total = isd ? row_total[threadIdx.x] : 0u;
  const uint32_t my_table = isd ? table[threadIdx.x * tiles + blockIdx.x] : 0u;

  uint32_t key[R];
  int32_t val[R];
  const int wbase = base + w * R * 64;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = wbase + r * 64 + lane;
    key[r] = i < n ? kin[i] : 0u;
    val[r] = i < n ? vin[i] : 0;
  }
  const uint32_t dmask = (1u << nbits) - 1u;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  __syncthreads();
  // pass 1: digit counts of this wave's sub-tile; pre[r]: keys of the same
  // digit before this one in the wave's sub-tile
  uint32_t pre[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool valid = wbase + r * 64 + lane < n;
    const uint32_t dg = (key[r] >> shift) & dmask;
    const uint64_t m = match_digit(dg, valid, nbits);
    const uint32_t old = valid ? wave_hist[w][dg] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (valid && (m & lt) == 0) wave_hist[w][dg] = old + (uint32_t)__popcll(m);
    __builtin_amdgcn_wave_barrier();
    pre[r] = old + (uint32_t)__popcll(m & lt);
  }
  __syncthreads();
  // tile-local digit starts (exclusive scan of the tile's digit counts), each
  // wave's first slot per digit, and the tile's output position per digit
  {
    const int d = threadIdx.x;
    uint32_t tot = 0;
    if (isd)
      for (int q = 0; q < W; ++q) tot += wave_hist[q][d];
    int total;
    const uint32_t start = (uint32_t)block_scan_u32(tot, &total);
    __syncthreads();  // block_scan_u32's LDS is reused below
    // digit bases: exclusive scan of the row totals (every workgroup does it)
    const uint32_t dbase = (uint32_t)block_scan_u32(my_row_total, &total);
    if (isd) {
      local_start[d] = start;
      gstart[d] = dbase + my_table;
      uint32_t run = start;
      for (int q = 0; q < W; ++q) {
        const uint32_t c = wave_hist[q][d];
        wave_hist[q][d] = run;
        run += c;
      }
    }
  }
  __syncthreads();
  // pass 2: each key to its slot in the tile's sorted order, in LDS
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool valid = wbase + r * 64 + lane < n;
    const uint32_t dg = (key[r] >> shift) & dmask;
    if (valid) {
      const uint32_t pos = wave_hist[w][dg] + pre[r];
      skey[pos] = key[r];
      sval[pos] = val[r];
    }
  }
  __syncthreads();
  // pass 3: the tile in sorted order, each digit's run to its output range
  // (consecutive slots go to consecutive addresses)
  for (int i = threadIdx.x; i < tile_n; i += kThreads) {
    const uint32_t k = skey[i];
    const uint32_t dg = (k >> shift) & dmask;
    const uint32_t o = gstart[dg] + ((uint32_t)i - local_start[dg]);
    kout[o] = k;
    vout[o] = sval[i];
  }
}

}  // namespace

int RadixSort::init(int n) {
  release();
  max_n = n < 1 ? 1 : n;
  max_tiles = (max_n + kTile - 1) / kTile;
  if (hipMalloc(&keys_tmp, sizeof(uint32_t) * max_n) != hipSuccess ||
      hipMalloc(&vals_tmp, sizeof(int32_t) * max_n) != hipSuccess ||
      hipMalloc(&table, sizeof(uint32_t) * kBins * max_tiles) != hipSuccess ||
      hipMalloc(&bases, sizeof(uint32_t) * kBins) != hipSuccess) {
    (void)hipGetLastError();
    release();
    return fail(NBX_ERR_NOMEM, "radix sort buffers for %d keys", n);
  }
  return NBX_OK;
}

void RadixSort::release() {
  if (keys_tmp) (void)hipFree(keys_tmp);
  if (vals_tmp) (void)hipFree(vals_tmp);
  if (table) (void)hipFree(table);
  if (bases) (void)hipFree(bases);
  bases = nullptr;
  keys_tmp = nullptr;
  vals_tmp = nullptr;
  table = nullptr;
}

int RadixSort::sort(uint32_t* keys, int32_t* vals, const int32_t* n_dev, int key_bits,
                    hipStream_t s) {
  if (key_bits <= 0) return NBX_OK;
  uint32_t* ka = keys;
  int32_t* va = vals;
  uint32_t* kb = keys_tmp;
  int32_t* vb = vals_tmp;
  int passes = 0;
  for (int shift = 0; shift < key_bits; shift += 8) {
    const int nbits = key_bits - shift < 8 ? key_bits - shift : 8;
    hipLaunchKernelGGL(radix_count, dim3(max_tiles), dim3(kThreads), 0, s, ka, n_dev, shift, table,
                       max_tiles);
    hipLaunchKernelGGL(radix_scan_rows, dim3(kBins), dim3(256), 0, s, table, max_tiles, bases);
    hipLaunchKernelGGL(radix_scatter, dim3(max_tiles), dim3(kThreads), 0, s, ka, va, kb, vb, table,
                       bases, n_dev, shift, nbits, max_tiles);
    NBX_LAUNCHED();
    uint32_t* tk = ka;
    ka = kb;
    kb = tk;
    int32_t* tv = va;
    va = vb;
    vb = tv;
    ++passes;
  }
  if (passes & 1) {  // result sits in the scratch buffers: copy back
    NBX_HIP(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * max_n, hipMemcpyDeviceToDevice, s));
    NBX_HIP(hipMemcpyAsync(vals, va, sizeof(int32_t) * max_n, hipMemcpyDeviceToDevice, s));
  }
  return NBX_OK;
}

}  // namespace nbx
