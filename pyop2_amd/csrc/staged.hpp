// LDS-staged gather / compute / scatter for Dat INC loops (OP2HIP_STAGED).
//
// One workgroup per tile of elements (op2hip_plan_tiles: compact regions,
// per staged map the tile's distinct nodes, those no other tile references
// first).  The tile's node values of every READ argument are gathered once
// into LDS, each thread runs the element kernel of one element on values
// read from LDS through its 16-bit local node ids, and adds its increments
// into LDS accumulators (one per tile node and component).  The tile then
// writes each node once: a plain read-modify-write for the nodes only this
// tile has, a device-scope atomic add for the nodes it shares.  Compared
// with one atomic per element and node this cuts device atomics by the
// number of elements around a node, and every gather becomes an LDS read.
//
// Elem provides
//   using T;  static constexpr int kMaps (1 or 2);
//   static constexpr int kAr0, kIn0, kOut0, kAr1, kIn1, kOut1;
//     arity, staged values per node (READ arguments on that map, in the
//     order stage() writes them) and increments per node (INC arguments),
//     for map 0 and map 1 (kAr1 = 0 when kMaps == 1);
//   template <int M> void stage(int32_t node, T *dst) const;   // kInM values
//   template <int M> T *out() const;                             // INC target, kOutM per node
//   void element(int32_t e, const T (&x0)[A0][I0], const T (&x1)[A1][I1],
//                T (&y0)[A0][O0], T (&y1)[A1][O1]) const;     // A/I/O >= 1
//   or, with kLdsElement = true, element_lds(e, in0, l0, in1, l1, y0, y1)
//   reading the staged records from LDS itself (see LdsElementOf).
#pragma once

#include "launch.hpp"

namespace op2hip {

struct TilePlan {
  int32_t ntiles;
  const int32_t *tile_ptr;
  const int32_t *tile_elems;
  const int32_t *node_ptr[2];
  const int32_t *nodes[2];
  const int32_t *nexcl[2];
  const uint16_t *local[2];
  int32_t max_nodes[2];
  // shared-node partials (null shared_off[m]: device atomics)
  const int32_t *shared_off[2];
  double *partials[2];
  // coloured tiles: the launch runs the tiles tile_list[0..gridDim.x), which
  // share no node, so every node of a tile is written like an exclusive one
  // (null: tile blockIdx.x of all ntiles, shared nodes as above)
  const int32_t *tile_list;
  bool all_exclusive;
};

constexpr int at_least_1(int v) { return v > 0 ? v : 1; }

// LDS: per map, staged inputs (T) then accumulators (double, 8-byte aligned).
template <class Elem>
struct TileLayout {
  using T = typename Elem::T;
  __host__ __device__ static size_t align8(size_t b) { return (b + 7) & ~(size_t)7; }
  __host__ __device__ static size_t in_at(int m, const int32_t *mx) {
    return m == 0 ? 0 : acc_at(0, mx) + (size_t)mx[0] * Elem::kOut0 * sizeof(double);
  }
  __host__ __device__ static size_t acc_at(int m, const int32_t *mx) {
    return align8(in_at(m, mx) + (size_t)mx[m] * (m == 0 ? Elem::kIn0 : Elem::kIn1) * sizeof(T));
  }
  static size_t bytes(const int32_t *mx) {
    return Elem::kMaps == 1 ? acc_at(0, mx) + (size_t)mx[0] * Elem::kOut0 * sizeof(double)
                            : acc_at(1, mx) + (size_t)mx[1] * Elem::kOut1 * sizeof(double);
  }
};

// kStageSlots: node slots per thread and staging round; the ids of a
// thread's first round stay in registers from staging to flush.  When a
// tile has at most kStageSlots * blockDim nodes the tile stages in one
// round: every node id of the thread, then every value, one dependent wait.

// Stage map M of tile b: inputs into LDS, and the accumulators start from the
// old output values of the nodes only this tile has (no other tile of the
// launch reads or writes them, so reading them here instead of at the flush
// changes nothing) and from zero for shared nodes.  All loads of a thread's
// node slots are issued together.  ids[] keeps the node ids of the thread's
// first kStageSlots slots for the flush.
template <class Elem, int M, int kStageSlots>
__device__ __forceinline__ void stage_map(const Elem &elem, const TilePlan &P, int32_t b,
                                          typename Elem::T *in, double *acc,
                                          int32_t (&ids)[kStageSlots]) {
  using T = typename Elem::T;
  constexpr int kIn = M == 0 ? Elem::kIn0 : Elem::kIn1;
  constexpr int kOut = M == 0 ? Elem::kOut0 : Elem::kOut1;
  const int32_t n0 = P.node_ptr[M][b], nn = P.node_ptr[M][b + 1] - n0;
  const int32_t ne = kOut > 0 ? (P.all_exclusive ? nn : P.nexcl[M][b]) : 0;
  const int32_t bd = (int32_t)blockDim.x;
  const T *out = nullptr;
  if constexpr (kOut > 0) out = elem.template out<M>();
  for (int32_t k0 = threadIdx.x; k0 < nn; k0 += kStageSlots * bd) {
    int32_t id[kStageSlots];
#pragma unroll
    for (int u = 0; u < kStageSlots; ++u) {
      const int32_t k = k0 + u * bd;
      id[u] = k < nn ? P.nodes[M][n0 + k] : 0;
    }
    if (k0 == (int32_t)threadIdx.x) {
#pragma unroll
      for (int u = 0; u < kStageSlots; ++u) ids[u] = id[u];
    }
#pragma unroll
    for (int u = 0; u < kStageSlots; ++u) {
      const int32_t k = k0 + u * bd;
      if (k >= nn) continue;
      if constexpr (kIn > 0) elem.template stage<M>(id[u], in + (size_t)k * kIn);
#pragma unroll
      for (int a = 0; a < kOut; ++a)
        acc[(size_t)k * kOut + a] = k < ne ? double(out[(size_t)id[u] * kOut + a]) : 0.0;
    }
  }
}

// Flush map M of tile b: exclusive nodes take their accumulator (which
// started from the old value); shared nodes write it to their partial slot
// (contiguous per tile) when the plan has partials, else add it with a
// device atomic.  No load waits.
template <class Elem, int M, int kStageSlots>
__device__ __forceinline__ void flush_map(const Elem &elem, const TilePlan &P, int32_t b,
                                          const double *acc, const int32_t (&ids)[kStageSlots]) {
  using T = typename Elem::T;
  constexpr int kOut = M == 0 ? Elem::kOut0 : Elem::kOut1;
  if constexpr (kOut > 0) {
    const int32_t n0 = P.node_ptr[M][b], nn = P.node_ptr[M][b + 1] - n0;
    const int32_t ne = P.all_exclusive ? nn : P.nexcl[M][b];
    const int32_t bd = (int32_t)blockDim.x;
    T *out = elem.template out<M>();
    double *part = nullptr;
    if (P.shared_off[M] && !P.all_exclusive)
      part = P.partials[M] + ((size_t)P.shared_off[M][b] - ne) * kOut;
    for (int32_t k0 = threadIdx.x; k0 < nn; k0 += kStageSlots * bd) {
#pragma unroll
      for (int u = 0; u < kStageSlots; ++u) {
        const int32_t k = k0 + u * bd;
        if (k >= nn) continue;
        if (k >= ne && part) {
#pragma unroll
          for (int a = 0; a < kOut; ++a) part[(size_t)k * kOut + a] = acc[(size_t)k * kOut + a];
          continue;
        }
        const int32_t id = k0 == (int32_t)threadIdx.x ? ids[u] : P.nodes[M][n0 + k];
        T *p = out + (size_t)id * kOut;
#pragma unroll
        for (int a = 0; a < kOut; ++a) {
          const T v = T(acc[(size_t)k * kOut + a]);
          if (k < ne) p[a] = v;
          else atomic_add(p + a, v);
        }
      }
    }
  }
}

// Second pass of a partials plan: every shared node adds its partials, in
// tile order, to the Dat (one thread per node; the order of the additions
// does not depend on scheduling).
template <typename T, int kOut>
__global__ void __launch_bounds__(256)
staged_reduce_kernel(int32_t nshared, const int32_t *__restrict__ nodes,
                     const int32_t *__restrict__ ptr, const int32_t *__restrict__ slots,
                     const double *__restrict__ partials, T *__restrict__ out) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nshared) return;
  double sum[kOut];
#pragma unroll
  for (int a = 0; a < kOut; ++a) sum[a] = 0.0;
  for (int32_t j = ptr[s]; j < ptr[s + 1]; ++j) {
    const double *p = partials + (size_t)slots[j] * kOut;
#pragma unroll
    for (int a = 0; a < kOut; ++a) sum[a] += p[a];
  }
  T *o = out + (size_t)nodes[s] * kOut;
#pragma unroll
  for (int a = 0; a < kOut; ++a) o[a] = T(double(o[a]) + sum[a]);
}

// Whether an element reads its staged values from LDS itself
// (Elem::kLdsElement, with element_lds(e, in0, l0, in1, l1, y0, y1): in0 and
// in1 the tile's staged records in LDS, l0 and l1 the element's local node
// ids).  An element whose inputs each feed several terms keeps fewer values
// live by reading them where they are used than by taking them as arrays.
template <class Elem, class = void>
struct LdsElementOf {
  static constexpr bool value = false;
};
template <class Elem>
struct LdsElementOf<Elem, std::void_t<decltype(Elem::kLdsElement)>> {
  static constexpr bool value = Elem::kLdsElement;
};

// Tile b of a staged launch (the workgroup's share, or one work item of a
// batched launch).  The local node ids of a thread's first element are
// loaded before the staging barrier.
template <class Elem, int kStageSlots = 2>
__device__ __forceinline__ void staged_tile(const Elem &elem, const TilePlan &P, int32_t b) {
  using T = typename Elem::T;
  using L = TileLayout<Elem>;
  constexpr int A0 = Elem::kAr0, I0 = at_least_1(Elem::kIn0), O0 = at_least_1(Elem::kOut0);
  constexpr int A1 = at_least_1(Elem::kAr1), I1 = at_least_1(Elem::kIn1),
                O1 = at_least_1(Elem::kOut1);
  extern __shared__ __align__(16) unsigned char smem[];
  T *in0 = reinterpret_cast<T *>(smem + L::in_at(0, P.max_nodes));
  double *acc0 = reinterpret_cast<double *>(smem + L::acc_at(0, P.max_nodes));
  T *in1 = reinterpret_cast<T *>(smem + (Elem::kMaps > 1 ? L::in_at(1, P.max_nodes) : 0));
  double *acc1 =
      reinterpret_cast<double *>(smem + (Elem::kMaps > 1 ? L::acc_at(1, P.max_nodes) : 0));

  const int32_t p0 = P.tile_ptr[b], ne = P.tile_ptr[b + 1] - p0;
  // Element of the tile that thread slot t (round t / blockDim) computes:
  // the lanes of a wave are spread over the whole tile (stride rounds *
  // waves), each lane's elements of successive rounds adjacent.  A tile's
  // elements are in breadth-first order, so neighbours share nodes; spread
  // lanes share fewer, and a wave's LDS adds collide less (float64 P2
  // residual at 64^3: 107 against 125 us with lane l on element l).  The
  // launchers take block sizes that are multiples of 64.
  const int32_t bdim = (int32_t)blockDim.x, nwaves = bdim / 64;
  const int32_t rounds = (ne + bdim - 1) / bdim;
  auto pos_of = [&](int32_t t) {
    const int32_t tt = t % bdim, r = t / bdim;
    return (tt % 64) * rounds * nwaves + (tt / 64) * rounds + r;
  };
  uint16_t l0[A0], l1[A1];
  int32_t elem_id = 0;
  auto load_locals = [&](int32_t pos) {
#pragma unroll
    for (int i = 0; i < A0; ++i) l0[i] = P.local[0][(size_t)pos * A0 + i];
    if constexpr (Elem::kMaps > 1) {
#pragma unroll
      for (int i = 0; i < A1; ++i) l1[i] = P.local[1][(size_t)pos * A1 + i];
    }
    elem_id = P.tile_elems[pos];
  };
  if (pos_of(threadIdx.x) < ne) load_locals(p0 + pos_of(threadIdx.x));

  int32_t ids0[kStageSlots], ids1[kStageSlots];
  stage_map<Elem, 0, kStageSlots>(elem, P, b, in0, acc0, ids0);
  if constexpr (Elem::kMaps > 1) stage_map<Elem, 1, kStageSlots>(elem, P, b, in1, acc1, ids1);
  __syncthreads();

  for (int32_t t = threadIdx.x; t < rounds * bdim; t += bdim) {
    const int32_t pos = pos_of(t);
    if (pos >= ne) continue;
    if (t != (int32_t)threadIdx.x) load_locals(p0 + pos);
    T y0[A0][O0], y1[A1][O1];
    if constexpr (LdsElementOf<Elem>::value) {
      elem.element_lds(elem_id, in0, l0, in1, l1, y0, y1);
    } else {
      T x0[A0][I0], x1[A1][I1];
#pragma unroll
      for (int i = 0; i < A0; ++i)
#pragma unroll
        for (int a = 0; a < I0; ++a) x0[i][a] = Elem::kIn0 > 0 ? in0[(size_t)l0[i] * I0 + a] : T(0);
#pragma unroll
      for (int i = 0; i < A1; ++i)
#pragma unroll
        for (int a = 0; a < I1; ++a)
          x1[i][a] = (Elem::kMaps > 1 && Elem::kIn1 > 0) ? in1[(size_t)l1[i] * I1 + a] : T(0);
      elem.element(elem_id, x0, x1, y0, y1);
    }
    if constexpr (Elem::kOut0 > 0) {
#pragma unroll
      for (int i = 0; i < A0; ++i)
#pragma unroll
        for (int a = 0; a < O0; ++a) atomic_add(&acc0[(size_t)l0[i] * O0 + a], double(y0[i][a]));
    }
    if constexpr (Elem::kMaps > 1 && Elem::kOut1 > 0) {
#pragma unroll
      for (int i = 0; i < A1; ++i)
#pragma unroll
        for (int a = 0; a < O1; ++a) atomic_add(&acc1[(size_t)l1[i] * O1 + a], double(y1[i][a]));
    }
  }
  __syncthreads();

  flush_map<Elem, 0, kStageSlots>(elem, P, b, acc0, ids0);
  if constexpr (Elem::kMaps > 1) flush_map<Elem, 1, kStageSlots>(elem, P, b, acc1, ids1);
}

// Waves per SIMD an element asks the compiler to fit (Elem::kWavesPerEU:
// the register budget follows, 512 / waves), default 1 (no constraint).
template <class Elem, class = void>
struct WavesPerEUOf {
  static constexpr int value = 1;
};
template <class Elem>
struct WavesPerEUOf<Elem, std::void_t<decltype(Elem::kWavesPerEU)>> {
  static constexpr int value = Elem::kWavesPerEU;
};

template <class Elem, int kStageSlots>
__global__ void __launch_bounds__(256, WavesPerEUOf<Elem>::value)
staged_kernel(Elem elem, TilePlan P) {
  staged_tile<Elem, kStageSlots>(
      elem, P, P.tile_list ? P.tile_list[blockIdx.x] : xcd_contiguous(blockIdx.x, P.ntiles));
}

template <class Elem>
int launch_staged(int32_t start, int32_t end, const Elem &elem, const op2hip_exec *ex,
                  hipStream_t s) {
  OP2HIP_REQUIRE(ex && ex->tile_ptr && ex->tile_elems && ex->ntiles >= 0,
                 "staged launch without a tile plan");
  OP2HIP_REQUIRE(!ex->subset, "staged launches do not iterate Subsets");
  for (int m = 0; m < Elem::kMaps; ++m)
    OP2HIP_REQUIRE(ex->tile_node_ptr[m] && ex->tile_nodes[m] && ex->tile_nexcl[m] &&
                       ex->elem_local[m],
                   "staged launch: tile plan lacks map %d", m);
  if (end <= start || ex->ntiles == 0) return OP2HIP_OK;
  for (int m = 0; m < Elem::kMaps; ++m)
    OP2HIP_REQUIRE(!ex->tile_shared_off[m] ||
                       (ex->partials[m] && ex->shared_nodes[m] && ex->shared_ptr[m] &&
                        ex->shared_slots[m] && ex->nshared[m] >= 0),
                   "staged launch: map %d has a partial-slot plan without its arrays", m);
  TilePlan P{ex->ntiles, ex->tile_ptr, ex->tile_elems,
             {ex->tile_node_ptr[0], ex->tile_node_ptr[1]},
             {ex->tile_nodes[0], ex->tile_nodes[1]},
             {ex->tile_nexcl[0], ex->tile_nexcl[1]},
             {ex->elem_local[0], ex->elem_local[1]},
             {ex->max_tile_nodes[0], Elem::kMaps > 1 ? ex->max_tile_nodes[1] : 0},
             {Elem::kOut0 > 0 ? ex->tile_shared_off[0] : nullptr,
              Elem::kMaps > 1 && Elem::kOut1 > 0 ? ex->tile_shared_off[1] : nullptr},
             {ex->partials[0], Elem::kMaps > 1 ? ex->partials[1] : nullptr},
             nullptr, false};
  const size_t lds = TileLayout<Elem>::bytes(P.max_nodes);
  OP2HIP_REQUIRE(lds <= 64 * 1024, "tile needs %zu bytes of LDS (max 65536)", lds);
  const int bs = block_size_of(ex, 256);
  OP2HIP_REQUIRE(bs > 0 && bs <= 256 && bs % 64 == 0, "staged block size %d", bs);
  // tiles whose nodes take at most four per thread stage in one round
  const int most = std::max(P.max_nodes[0], Elem::kMaps > 1 ? P.max_nodes[1] : 0);
  auto run = [&](int32_t grid) {
    if (most <= 4 * bs)
      staged_kernel<Elem, 4><<<grid, bs, lds, s>>>(elem, P);
    else
      staged_kernel<Elem, 2><<<grid, bs, lds, s>>>(elem, P);
  };
  if (ex->ncolours > 0) {
    // coloured tiles: one launch per colour; the tiles of a colour share no
    // node, so no partials, atomics or second pass
    OP2HIP_REQUIRE(ex->colour_elems && ex->colour_offsets,
                   "coloured staged launch without its tile colouring");
    OP2HIP_REQUIRE(ex->colour_offsets[0] == 0 && ex->colour_offsets[ex->ncolours] == ex->ntiles,
                   "tile colouring covers %d of %d tiles", ex->colour_offsets[ex->ncolours],
                   ex->ntiles);
    P.all_exclusive = true;
    P.shared_off[0] = P.shared_off[1] = nullptr;
    for (int32_t c = 0; c < ex->ncolours; ++c) {
      const int32_t lo = ex->colour_offsets[c], hi = ex->colour_offsets[c + 1];
      OP2HIP_REQUIRE(hi >= lo, "tile colouring offsets decrease at colour %d", c);
      if (hi == lo) continue;
      P.tile_list = ex->colour_elems + lo;
      run(hi - lo);
    }
    OP2HIP_LAUNCHED();
    return OP2HIP_OK;
  }
  run(ex->ntiles);
  OP2HIP_LAUNCHED();
  if constexpr (Elem::kOut0 > 0) {
    if (P.shared_off[0] && ex->nshared[0] > 0)
      staged_reduce_kernel<typename Elem::T, Elem::kOut0><<<(ex->nshared[0] + 255) / 256, 256, 0, s>>>(
          ex->nshared[0], ex->shared_nodes[0], ex->shared_ptr[0], ex->shared_slots[0],
          ex->partials[0], elem.template out<0>());
  }
  if constexpr (Elem::kMaps > 1 && Elem::kOut1 > 0) {
    if (P.shared_off[1] && ex->nshared[1] > 0)
      staged_reduce_kernel<typename Elem::T, Elem::kOut1><<<(ex->nshared[1] + 255) / 256, 256, 0, s>>>(
          ex->nshared[1], ex->shared_nodes[1], ex->shared_ptr[1], ex->shared_slots[1],
          ex->partials[1], elem.template out<1>());
  }
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // namespace op2hip
