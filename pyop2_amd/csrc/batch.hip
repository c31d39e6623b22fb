// Batched P1 Laplace assembly + residual over many independent meshes in one
// persistent launch.
//
// Work items are (mesh, row batch of the mesh's CSR assembly) and (mesh, tile
// of the mesh's staged residual); the grid is sized to what the device holds
// at once and each workgroup walks the item list with a grid stride, so one
// launch (one node of a HIP graph) serves every mesh whatever its size.  An
// item runs exactly the code of the single-mesh launch: slots_batch for the
// slot-table row gather, staged_tile for the LDS-staged residual.
#include "p1.hpp"

namespace op2hip {

__global__ void __launch_bounds__(256)
batched_p1_kernel(const op2hip_batch_mesh *__restrict__ meshes, const int2 *__restrict__ items,
                  int32_t nitems) {
  for (int32_t w = blockIdx.x; w < nitems; w += gridDim.x) {
    const int2 it = items[w];
    const op2hip_batch_mesh &M = meshes[it.x];
    if ((it.y & OP2HIP_BATCH_TILE) == 0) {
      const BatchDesc d = static_cast<const BatchDesc *>(M.batch_desc)[it.y];
      slots_batch<P1MatrixRow, 8>(d, 0, M.ncells, P1MatrixRow{M.X, M.c2v}, M.A,
                                 static_cast<const SlotEntry<3> *>(M.slot_table),
                                 static_cast<double *>(nullptr));
    } else {
      const TilePlan P{M.ntiles, M.tile_ptr, M.tile_elems,
                       {M.tile_node_ptr, nullptr}, {M.tile_nodes, nullptr},
                       {M.tile_nexcl, nullptr}, {M.elem_local, nullptr},
                       {M.max_tile_nodes, 0}, {nullptr, nullptr}, {nullptr, nullptr},
                       nullptr, false};
      staged_tile(P1ResidualStaged{M.R, M.X, M.u}, P, it.y & ~OP2HIP_BATCH_TILE);
    }
    __syncthreads();  // the next item reuses the LDS
  }
}

}  // namespace op2hip

using namespace op2hip;

extern "C" {

int op2hip_batched_p1_laplace(const op2hip_batch_mesh *meshes, const int32_t *items,
                              int32_t nitems, int32_t grid, int32_t max_batch_nnz,
                              int32_t max_tile_nodes, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(meshes && items && nitems >= 0 && grid >= 0 && max_batch_nnz >= 0 &&
                     max_tile_nodes >= 0,
                 "batched_p1_laplace: bad arguments");
  if (nitems == 0) return OP2HIP_OK;
  const int32_t mx[2] = {max_tile_nodes, 0};
  const size_t lds = std::max((size_t)max_batch_nnz * sizeof(double),
                              TileLayout<P1ResidualStaged>::bytes(mx));
  OP2HIP_REQUIRE(lds <= 64 * 1024, "batched launch needs %zu bytes of LDS (max 65536)", lds);
  const int32_t g = grid > 0 ? std::min(grid, nitems) : nitems;
  batched_p1_kernel<<<g, 256, lds, s>>>(meshes, reinterpret_cast<const int2 *>(items), nitems);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // extern "C"
