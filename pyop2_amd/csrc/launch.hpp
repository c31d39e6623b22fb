// Launch helpers shared by the element-loop entry points.
#pragma once

#include "common.hpp"

namespace op2hip {

// Calls launch(lo, hi, list, atomic) once per conflict-free launch of the
// iteration range [start, end):
//  * OP2HIP_ATOMIC (and no plan): one launch over [start, end), elements
//    indirected through ex->subset when iterating a Subset, atomic scatter;
//  * OP2HIP_COLOURED: one launch per colour over colour_elems[off[c],
//    off[c+1]), plain read-modify-write scatter (no two elements of a colour
//    share a target).
template <class F>
int for_each_launch(const op2hip_exec *ex, int32_t start, int32_t end, F &&launch) {
  if (end <= start) return OP2HIP_OK;
  const int strategy = ex ? ex->strategy : OP2HIP_ATOMIC;
  if (strategy == OP2HIP_COLOURED) {
    OP2HIP_REQUIRE(ex->colour_elems && ex->colour_offsets && ex->ncolours >= 0,
                   "coloured launch without a colour plan");
    OP2HIP_REQUIRE(ex->colour_offsets[0] == 0 &&
                       ex->colour_offsets[ex->ncolours] == end - start,
                   "colour plan covers %d elements, launch range has %d",
                   ex->colour_offsets[ex->ncolours], end - start);
    for (int c = 0; c < ex->ncolours; ++c) {
      const int32_t lo = ex->colour_offsets[c], hi = ex->colour_offsets[c + 1];
      if (hi > lo) {
        int rc = launch(lo, hi, ex->colour_elems, false);
        if (rc) return rc;
      }
    }
    return OP2HIP_OK;
  }
  OP2HIP_REQUIRE(strategy == OP2HIP_ATOMIC || strategy == OP2HIP_STAGED ||
                     strategy == OP2HIP_ROWGATHER || strategy == OP2HIP_ROWGATHER_THREAD ||
                     strategy == OP2HIP_ROWGATHER_SLOTS || strategy == OP2HIP_ROWGATHER_STAGED ||
                     strategy == OP2HIP_ROWGATHER_CELLS || strategy == OP2HIP_ROWGATHER_ROWLIST,
                 "unknown strategy %d", strategy);
  return launch(start, end, ex ? ex->subset : nullptr, true);
}

inline int block_size_of(const op2hip_exec *ex, int dflt) {
  return (ex && ex->block_size > 0) ? ex->block_size : dflt;
}

// Workgroup id -> batch id so that each XCD (workgroups are dispatched to
// the 8 XCDs round-robin) works on one contiguous range of batches and its
// L2 sees the shared rows' neighbourhood.
__device__ __forceinline__ int32_t xcd_contiguous(int32_t bid, int32_t nb) {
  constexpr int32_t kXcd = 8;
  const int32_t x = bid % kXcd, idx = bid / kXcd;
  const int32_t q = nb / kXcd, rem = nb % kXcd;
  return x * q + min(x, rem) + idx;
}

// Workgroup id -> batch id dealing chunks of kChunk consecutive batches to
// the XCDs in turn: the chip works through one window of 8 * kChunk batches
// at a time, each XCD on kChunk neighbours of it.  Workgroups of the last,
// partial window keep their own id.
template <int32_t kChunk>
__device__ __forceinline__ int32_t xcd_dealt(int32_t bid, int32_t nb) {
  constexpr int32_t kXcd = 8, kWindow = kXcd * kChunk;
  const int32_t w = bid / kWindow, r = bid % kWindow;
  return (w + 1) * kWindow <= nb ? w * kWindow + (r % kXcd) * kChunk + r / kXcd : bid;
}

}  // namespace op2hip
