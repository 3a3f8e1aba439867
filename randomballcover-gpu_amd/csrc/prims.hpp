// Data-movement primitives of librbc_hip: exclusive scan, stable counting sort by
// bucket, and work-item generation.  All run on a caller-given stream and use
// scratch memory owned by a Workspace.
#pragma once

#include <cstdint>
#include <vector>

#include "rbc_common.hpp"
#include "rbc_tile.hpp"

namespace rbc {

// Grow-only device scratch shared by the calls of one handle.
struct Workspace {
  std::vector<DevBuf<unsigned char>> bufs;
  // Returns slot `i` with at least `bytes` bytes.
  int get(size_t i, size_t bytes, void **out) {
    if (bufs.size() <= i) bufs.resize(i + 1);
    RBC_TRY(bufs[i].ensure(bytes));
    *out = bufs[i].get();
    return RBC_OK;
  }
  size_t bytes() const {
    size_t s = 0;
    for (auto &b : bufs) s += b.n;
    return s;
  }
};

// out[i] = sum_{j<i} in[i] for i <= n (out has n+1 entries; out[n] = total).
// in and out may alias.  Uses workspace slots [slot, slot + 2L), L = the number of levels
// (ceil(log_2048 n), at most 3 below 2^33 elements).
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, Workspace &ws, size_t slot,
                       hipStream_t st);

// Stable counting sort of items 0..n-1 by key[i] in [0, nbuckets): order[pos] = i with
// items of a bucket in ascending i; offsets[b] = first position of bucket b
// (nbuckets+1 entries; offsets[nbuckets] = number of items with a valid key -- items
// whose key is >= nbuckets are left out).  Uses workspace slots [slot, slot + 1 + 2L), L the
// levels of the scan over nbuckets x ceil(n / 4096) counts.
int counting_sort(const uint32_t *key, size_t n, uint32_t nbuckets, uint32_t *order,
                  uint32_t *offsets, Workspace &ws, size_t slot, hipStream_t st);

// For every bucket b with cnt_b = qoff[b+1]-qoff[b] queries, emits ceil(cnt_b/tile) items
// {x_begin = xoff[b], x_count = xlen[b], q_begin = qoff[b] + tile t, q_count, aux = b}.
// Returns the number of items on the host (one synchronisation).  Items are written
// bucket-major.  Uses workspace slots [slot, slot + 1 + 2L), L the levels of the scan over
// nbuckets counts, and items_buf.
int make_bucket_items(const uint32_t *qoff, const uint32_t *xoff, const uint32_t *xlen,
                      uint32_t nbuckets, DevBuf<ScanItem> &items_buf, uint32_t *nitems,
                      Workspace &ws, size_t slot, hipStream_t st, uint32_t tile = kTQ);

// Items covering queries [0, m) in tiles of 64 against points [0, n) split into
// segments of at most seg_len points; aux = segment index.
std::vector<ScanItem> grid_items(uint32_t m, uint32_t n, uint32_t seg_len);

}  // namespace rbc
