// Radix sort of binned records by (bin, footprint origin) with hipCUB: the
// records of a (layer, uv tile) bin that share a footprint position become
// consecutive, so the grid kernel adds each footprint's kernel-weighted sum
// to its tile once per run instead of once per record.
#include "hws_internal.h"

#include <hipcub/hipcub.hpp>

namespace hws {

hws_status sortRecords(hws_gridder* g, uint32_t*& keys, uint32_t*& keysAlt, uint2*& recs,
                       uint2*& recsAlt, size_t n, int endBit) {
  if (n < 2) return HWS_OK;
  Plan& pl = g->plan;
  hipcub::DoubleBuffer<uint32_t> k(keys, keysAlt);
  hipcub::DoubleBuffer<uint2> v(recs, recsAlt);
  size_t bytes = 0;
  hws_status st = checkHip(
      g, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, n, 0, endBit, g->stream),
      "record sort size");
  if (st != HWS_OK) return st;
  if (pl.sortTempCapacity < bytes) {
    if (pl.sortTemp) (void)hipFree(pl.sortTemp);
    pl.sortTemp = nullptr;
    pl.sortTempCapacity = 0;
    st = checkHip(g, hipMalloc(&pl.sortTemp, bytes), "record sort scratch");
    if (st != HWS_OK) return st;
    pl.sortTempCapacity = bytes;
  }
  st = checkHip(g,
                hipcub::DeviceRadixSort::SortPairs(pl.sortTemp, bytes, k, v, n, 0, endBit,
                                                   g->stream),
                "record sort");
  if (st != HWS_OK) return st;
  if (k.Current() != keys) std::swap(keys, keysAlt);
  if (v.Current() != recs) std::swap(recs, recsAlt);
  return HWS_OK;
}

// Row indices ordered by 32-bit key (the binning's row groups).
hws_status sortRows(hws_gridder* g, uint32_t*& keys, uint32_t*& keysAlt, uint32_t*& rows,
                    uint32_t*& rowsAlt, size_t n) {
  if (n < 2) return HWS_OK;
  Plan& pl = g->plan;
  hipcub::DoubleBuffer<uint32_t> k(keys, keysAlt);
  hipcub::DoubleBuffer<uint32_t> v(rows, rowsAlt);
  size_t bytes = 0;
  hws_status st = checkHip(
      g, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, n, 0, 32, g->stream),
      "row sort size");
  if (st != HWS_OK) return st;
  if (pl.sortTempCapacity < bytes) {
    if (pl.sortTemp) (void)hipFree(pl.sortTemp);
    pl.sortTemp = nullptr;
    pl.sortTempCapacity = 0;
    st = checkHip(g, hipMalloc(&pl.sortTemp, bytes), "row sort scratch");
    if (st != HWS_OK) return st;
    pl.sortTempCapacity = bytes;
  }
  st = checkHip(g,
                hipcub::DeviceRadixSort::SortPairs(pl.sortTemp, bytes, k, v, n, 0, 32,
                                                   g->stream),
                "row sort");
  if (st != HWS_OK) return st;
  if (k.Current() != keys) std::swap(keys, keysAlt);
  if (v.Current() != rows) std::swap(rows, rowsAlt);
  return HWS_OK;
}

}  // namespace hws
