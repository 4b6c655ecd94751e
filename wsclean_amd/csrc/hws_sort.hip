// Stable LSD radix sort of (32-bit key, value) pairs, 8 key bits per pass:
// the rows of a plan by their middle channel's bin (the binning's row
// groups), and binned records by (bin, footprint origin) so that the grid
// kernels for tile edges 24 and 32 add each footprint's kernel-weighted sum
// to its tile once per run of equal origins.
//
// Each pass: (1) every tile of kSortTile consecutive pairs histograms its
// digits in LDS and stores the counts digit-major, hist[d * nTiles + t];
// (2) one exclusive scan of that array gives every (digit, tile) its first
// output position (all smaller digits first, then the same digit's earlier
// tiles); (3) every tile reads its pairs again in index order, 256 at a time,
// ranks each pair among the pairs of its digit (a wave matches equal digits
// with eight ballots, the four waves' counts add up in LDS) and writes it to
// its position. Ranks follow index order, so the sort is stable and its
// output does not depend on scheduling.
#include "hws_internal.h"

namespace hws {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kDigits = 256;
constexpr int kSortWaves = kSortThreads / 64;

#define HWS_SORT_TRY(expr)       \
  do {                           \
    hws_status _s = (expr);      \
    if (_s != HWS_OK) return _s; \
  } while (0)

__global__ void __launch_bounds__(kSortThreads)
digitHistogramKernel(const uint32_t* __restrict__ keys, size_t n, int shift, uint32_t nTiles,
                     uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[kDigits];
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  const size_t base = size_t(blockIdx.x) * kSortTile;
#pragma unroll 4
  for (int r = 0; r < kSortRounds; ++r) {
    const size_t i = base + size_t(r) * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kDigits - 1)], 1u);
  }
  __syncthreads();
  hist[size_t(threadIdx.x) * nTiles + blockIdx.x] = cnt[threadIdx.x];
}

template <typename V>
__global__ void __launch_bounds__(kSortThreads)
digitScatterKernel(const uint32_t* __restrict__ keys, const V* __restrict__ vals, size_t n,
                   int shift, uint32_t nTiles, const uint32_t* __restrict__ offs,
                   uint32_t* __restrict__ keysOut, V* __restrict__ valsOut) {
  __shared__ uint32_t cursor[kDigits];               // next output position per digit
  __shared__ uint32_t waveCnt[kSortWaves][kDigits];  // this round's count per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  cursor[threadIdx.x] = offs[size_t(threadIdx.x) * nTiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < kSortWaves; ++w) waveCnt[w][threadIdx.x] = 0u;
  __syncthreads();
  const size_t base = size_t(blockIdx.x) * kSortTile;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int r = 0; r < kSortRounds; ++r) {
    const size_t i = base + size_t(r) * kSortThreads + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0u;
    V val{};
    if (valid) {
      key = keys[i];
      val = vals[i];
    }
    const uint32_t d = (key >> shift) & (kDigits - 1);
    // lanes of this wave with the same digit
    unsigned long long same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const uint32_t rank = uint32_t(__popcll(same & below));
    if (valid && rank == 0u) waveCnt[wave][d] = uint32_t(__popcll(same));
    __syncthreads();
    if (valid) {
      uint32_t pos = cursor[d] + rank;
      for (int w = 0; w < wave; ++w) pos += waveCnt[w][d];
      keysOut[pos] = key;
      valsOut[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0u;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
      add += waveCnt[w][threadIdx.x];
      waveCnt[w][threadIdx.x] = 0u;
    }
    cursor[threadIdx.x] += add;
    __syncthreads();
  }
}

// Scratch of one sort of n pairs: the histogram, its scan and the scan's
// block sums, in 32-bit words.
size_t sortScratchWords(size_t n) {
  const size_t nTiles = (n + kSortTile - 1) / kSortTile;
  const size_t h = nTiles * kDigits;
  return h + (h + 1) + scanScratchSize(h);
}

template <typename V>
hws_status radixSort(hws_gridder* g, uint32_t*& keys, uint32_t*& keysAlt, V*& vals,
                     V*& valsAlt, size_t n, int endBit, const char* what) {
  if (n < 2 || endBit <= 0) return HWS_OK;
  if (n > size_t(0xFFFFFFFFu))
    return setLastError(g, HWS_ERR_UNSUPPORTED, "radix sort of more than 2^32 - 1 pairs");
  Plan& pl = g->plan;
  const size_t bytes = sortScratchWords(n) * sizeof(uint32_t);
  if (pl.sortTempCapacity < bytes) {
    if (pl.sortTemp) (void)hipFree(pl.sortTemp);
    pl.sortTemp = nullptr;
    pl.sortTempCapacity = 0;
    HWS_SORT_TRY(checkHip(g, hipMalloc(&pl.sortTemp, bytes), what));
    pl.sortTempCapacity = bytes;
  }
  const uint32_t nTiles = uint32_t((n + kSortTile - 1) / kSortTile);
  const size_t h = size_t(nTiles) * kDigits;
  uint32_t* hist = static_cast<uint32_t*>(pl.sortTemp);
  uint32_t* offs = hist + h;
  uint32_t* scratch = offs + h + 1;
  for (int shift = 0; shift < endBit; shift += 8) {
    hipLaunchKernelGGL(digitHistogramKernel, dim3(nTiles), dim3(kSortThreads), 0, g->stream,
                       keys, n, shift, nTiles, hist);
    HWS_SORT_TRY(checkHip(g, hipGetLastError(), what));
    HWS_SORT_TRY(exclusiveScan(g, hist, offs, h, scratch));
    hipLaunchKernelGGL((digitScatterKernel<V>), dim3(nTiles), dim3(kSortThreads), 0, g->stream,
                       keys, vals, n, shift, nTiles, offs, keysAlt, valsAlt);
    HWS_SORT_TRY(checkHip(g, hipGetLastError(), what));
    std::swap(keys, keysAlt);
    std::swap(vals, valsAlt);
  }
  return HWS_OK;
}

}  // namespace

hws_status sortRecords(hws_gridder* g, uint32_t*& keys, uint32_t*& keysAlt, uint2*& recs,
                       uint2*& recsAlt, size_t n, int endBit) {
  return radixSort(g, keys, keysAlt, recs, recsAlt, n, endBit, "record sort");
}

hws_status sortRows(hws_gridder* g, uint32_t*& keys, uint32_t*& keysAlt, uint32_t*& rows,
                    uint32_t*& rowsAlt, size_t n) {
  return radixSort(g, keys, keysAlt, rows, rowsAlt, n, 32, "row sort");
}

}  // namespace hws
