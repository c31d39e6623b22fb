/* op2hip -- MI355X (gfx950) kernels for PyOP2's indirect par_loop hot path.
 *
 * C ABI: plain pointers and sizes, no torch or Python types.  Every pointer
 * documented as "device" must address HIP device memory of the current
 * device; "host" pointers are ordinary host memory.  Streams are hipStream_t
 * passed as void* (NULL = the legacy default stream).  Every entry point
 * returns 0 on success or an OP2HIP_ERR_* code; op2hip_last_error() returns
 * the message of the last failure on the calling thread.  Entry points never
 * free or reallocate caller buffers and never synchronise the stream unless
 * documented (functions returning a host-side count do).
 *
 * The element-loop entry points replace the ctypes function the reference
 * JIT-compiles per par_loop, ``wrap_<kernel>(int start, int end, void*...)``
 * (pyop2/global_kernel.py:322-331, argtypes :410-417, argument order
 * pyop2/codegen/builder.py:943-962 and pyop2/parloop.py:203-212): the
 * iteration range, then each par_loop argument's data pointer in call order,
 * then the unique maps in first-use order.  A Mat argument, a PETSc handle in
 * the reference (pyop2/types/mat.py:567-569), is an op2hip_mat descriptor
 * here.  Two trailing arguments are added: the execution plan and the stream.
 */
#ifndef OP2HIP_H
#define OP2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OP2HIP_ABI_VERSION 2

enum op2hip_status {
  OP2HIP_OK = 0,
  OP2HIP_ERR_HIP = 1,        /* a HIP runtime call failed */
  OP2HIP_ERR_ARG = 2,        /* invalid argument (sizes, NULL, unsupported) */
  OP2HIP_ERR_WORKSPACE = 3,  /* workspace too small */
  OP2HIP_ERR_OVERFLOW = 4,   /* a count does not fit int32 (PETSc int) */
  OP2HIP_ERR_JIT = 5,        /* runtime compilation or module load failed */
  OP2HIP_ERR_PATTERN = 6     /* Mat insertion outside the sparsity pattern */
};

/* Conflict-resolution strategies for indirect INC (and Mat) arguments. */
enum op2hip_strategy {
  OP2HIP_ATOMIC = 0,     /* one thread per element, device-scope atomics */
  OP2HIP_COLOURED = 1,   /* greedy element colouring, one launch per colour */
  OP2HIP_STAGED = 2,     /* LDS-staged tiles, one atomic per tile node */
  OP2HIP_ROWGATHER = 3,  /* Mat only: owner-computes block rows, no atomics;
                            one lane per (row incidence, column block) */
  OP2HIP_ROWGATHER_THREAD = 4, /* as ROWGATHER, one thread per row incidence */
  OP2HIP_ROWGATHER_SLOTS = 5,  /* as ROWGATHER_THREAD, with a precomputed slot
                                  table (no pattern search in the kernel) */
  OP2HIP_ROWGATHER_STAGED = 6, /* slot-table plan; element block rows staged in
                                  LDS, each output block summed by one thread
                                  (no atomics, deterministic) */
  OP2HIP_ROWGATHER_CELLS = 7,  /* row batches; each element's geometry staged in
                                  LDS once per batch, each output block
                                  evaluated and summed by one thread
                                  (op2hip_plan_cell_gather; no atomics) */
  OP2HIP_ROWGATHER_ROWLIST = 8 /* as ROWGATHER_SLOTS over batches of rows taken
                                  in an element-local order instead of row-id
                                  order (op2hip_plan_row_list): the gathers of
                                  a batch stay in cache when the row numbering
                                  is scattered */
};

/* Execution plan of one launch.  Host struct; arrays as annotated. */
typedef struct op2hip_exec {
  int32_t strategy;            /* enum op2hip_strategy */
  int32_t block_size;          /* threads per workgroup (0: kernel default) */
  /* OP2HIP_COLOURED: elements of [start,end) sorted by colour.
     OP2HIP_STAGED with ncolours > 0 (coloured tiles, the plan's two-level
     colouring of doc/sphinx/source/plan.rst:46-51): colour_elems lists the
     tile ids 0..ntiles-1 sorted by colour, tiles of one colour sharing no
     node of a map with increments; each colour runs as one launch and every
     node of a tile is written with a plain read-modify-write (no partials,
     no atomics, no second pass; tile_shared_off is ignored). */
  const int32_t *colour_elems;     /* device, end-start entries (tiles: ntiles) */
  const int32_t *colour_offsets;   /* host, ncolours+1 offsets into colour_elems */
  int32_t ncolours;
  /* OP2HIP_STAGED: tiles of the elements of [start,end) (op2hip_plan_tiles);
     per staged map m (in the order the entry point documents): */
  int32_t ntiles;
  const int32_t *tile_ptr;          /* device, ntiles+1 offsets into tile_elems */
  const int32_t *tile_elems;        /* device, end-start element ids by tile */
  const int32_t *tile_node_ptr[2];  /* device, ntiles+1 offsets into tile_nodes[m] */
  const int32_t *tile_nodes[2];     /* device, node ids of each tile, exclusive first */
  const int32_t *tile_nexcl[2];     /* device, per tile: nodes no other tile has */
  const uint16_t *elem_local[2];    /* device, (end-start) x arity: local node ids */
  int32_t max_tile_nodes[2];        /* largest node count of a tile (sizes LDS) */
  /* Optional, per staged map with increments: the nodes a tile shares with
     other tiles go through partial sums instead of device atomics
     (op2hip_plan_tile_reduce).  Each tile writes its shared nodes'
     accumulators to partials; a second pass adds, for every shared node, its
     partials in tile order to the Dat.  NULL tile_shared_off[m]: device
     atomic adds. */
  const int32_t *tile_shared_off[2]; /* device, ntiles+1: first partial slot of each tile */
  int32_t nshared[2];                /* distinct shared nodes */
  const int32_t *shared_nodes[2];    /* device, nshared node ids */
  const int32_t *shared_ptr[2];      /* device, nshared+1 offsets into shared_slots */
  const int32_t *shared_slots[2];    /* device, partial slots of each shared node, tile order */
  double *partials[2];               /* device workspace, tile_shared_off[m][ntiles] x increments per node */
  /* OP2HIP_ROWGATHER: batches of consecutive block rows */
  int32_t nbatches;
  const int32_t *batch_rows;       /* device, nbatches+1 row boundaries */
  int32_t batch_max_nnz;           /* max blocks of a batch (sizes LDS) */
  int32_t batch_max_inc;           /* max incidences of a batch */
  int32_t batch_max_rows;          /* max rows of a batch */
  const void *batch_desc;          /* device, per batch {r0, nr, z0, nz, i0, ni, 0, 0} (int32) */
  const void *slot_table;          /* device, per row incidence: op2hip_slot_table layout */
  /* OP2HIP_ROWGATHER_STAGED: gather table (op2hip_gather_table_fill) */
  const uint8_t *gather_count;     /* device, per Mat block: element blocks added to it */
  const uint16_t *gather_list;     /* device, per batch: element block positions */
  /* OP2HIP_ROWGATHER_CELLS (op2hip_plan_cell_gather; gather_count as above) */
  const int32_t *cell_ptr;         /* device, nbatches+1 offsets into cells */
  const int32_t *cells;            /* device, the distinct elements of each batch */
  const uint16_t *block_order;     /* device, per Mat block: batch blocks in thread order */
  const int32_t *contrib_ptr;      /* device, nbatches+1 offsets into contrib */
  const uint16_t *contrib;         /* device, (u << 8) | (i << 4) | j per element block */
  int32_t max_batch_cells;         /* largest cell count of a batch (<= 256) */
  int32_t max_batch_contrib;       /* largest contribution count of a batch */
  /* OP2HIP_ROWGATHER_ROWLIST (op2hip_plan_row_list; batch_desc / slot_table
     then index the batch-ordered incidences) */
  const int32_t *row_order;        /* device, (row, block offset in batch) in processing order */
  const int32_t *batch_ent;        /* device, per batch incidence: element * arity + i */
  const int32_t *row_base;         /* device, optional: rowptr[row] per row_order entry
                                      (NULL: read from the Mat's rowptr) */
  /* Subset iteration: element n = subset[i] for i in [start,end) (or NULL) */
  const int32_t *subset;           /* device */
  /* OP2HIP_ROWGATHER_ROWLIST, optional (arity-3 row and column maps): one
     word per batch incidence in place of batch_ent and slot_table -- bits
     0-31 the incidence (element * 3 + i), 32-47 the row's block offset in
     the batch, 48-62 the three column slots (5 bits each, 31 = missing) */
  const uint64_t *batch_packed;    /* device, or NULL */
} op2hip_exec;

/* Device CSR / BSR matrix on a fixed pattern (the reference's PETSc AIJ or
 * BAIJ, pyop2/types/mat.py:672-729).  values holds nnz blocks of
 * rdim x cdim entries, row-major within a block.  rdim = cdim = 1 is CSR. */
typedef struct op2hip_mat {
  const int32_t *rowptr;   /* device, nrows+1 */
  const int32_t *colidx;   /* device, nnz, sorted within each row */
  void *values;            /* device, nnz*rdim*cdim of the matrix dtype */
  int32_t nrows, ncols;    /* block rows / block columns */
  int32_t rdim, cdim;      /* block shape */
  int32_t *status;         /* device int32[1]; insertions that missed the pattern are counted here */
  /* Row-incidence plan (transpose of the row map) for OP2HIP_ROWGATHER */
  const int32_t *inc_ptr;  /* device, nrows+1 */
  const int32_t *inc_ent;  /* device, element*rarity + local row, sorted per row */
} op2hip_mat;

/* One (row map, column map) pair of a Sparsity (pyop2/types/mat.py:118-176). */
typedef struct op2hip_map_pair {
  const int32_t *rmap;     /* device, (>= iter_size) x rarity */
  const int32_t *cmap;     /* device, (>= iter_size) x carity */
  int32_t rarity, carity;
  int32_t iter_size;       /* owned size of the iteration set (sparsity.pyx:262) */
} op2hip_map_pair;

#define OP2HIP_MAX_PAIRS 8

/* ---- library ---------------------------------------------------------- */
int op2hip_abi_version(void);
const char *op2hip_last_error(void);
/* Selects `device` and checks it is gfx950-capable and reachable. */
int op2hip_init(int32_t device);

/* ---- prefix sums ------------------------------------------------------- */
/* out[0] = 0, out[i+1] = out[i] + in[i] for i < n (int32; OVERFLOW if the
 * total exceeds INT32_MAX).  Writes the total to *total_host (synchronises). */
int op2hip_scan_workspace(int64_t n, size_t *bytes);
int op2hip_exclusive_scan(int64_t n, const int32_t *in, int32_t *out,
                          int64_t *total_host, void *work, size_t work_bytes,
                          void *stream);

/* ---- map transpose (row incidences) ------------------------------------ */
/* For a map m (iter_size x arity, entries < 0 ignored) over [0, iter_size):
 * inc_ptr[r] .. inc_ptr[r+1] lists e*arity + i for every m[e][i] == r,
 * ascending.  count: fills inc_ptr and *total_host (synchronises);
 * fill: writes inc_ent (inc_ptr from count).  Replaces nothing in the
 * reference: the row-gather plan of the device assembly. */
int op2hip_transpose_workspace(int32_t nrows, size_t *bytes);
int op2hip_transpose_count(int32_t nrows, const int32_t *map, int32_t arity,
                           int32_t iter_size, int32_t *inc_ptr,
                           int64_t *total_host, void *work, size_t work_bytes,
                           void *stream);
int op2hip_transpose_fill(int32_t nrows, const int32_t *map, int32_t arity,
                          int32_t iter_size, const int32_t *inc_ptr,
                          int32_t *inc_ent, void *work, size_t work_bytes,
                          void *stream);

/* ---- sparsity (replaces pyop2/sparsity.pyx:121-284, build_sparsity and
 *      fill_with_zeros, non-extruded, one DataSet pair) ------------------ */
/* Block pattern of the union over `pairs` of rmap[e,:] x cmap[e,:] for
 * e < iter_size, plus (i, i) for i < min(nrows, ncols) when set_diag.
 * Step 1 (count): rowptr (device, nrows+1) and *nnz_host (synchronises).
 * Step 2 (fill): colidx (device, nnz), sorted unique per row.
 * Each pair's row incidences come from op2hip_transpose_* (inc_ptr/inc_ent
 * arrays, one per pair, device pointers passed in host arrays). */
int op2hip_sparsity_count(int32_t nrows, int32_t ncols,
                          const op2hip_map_pair *pairs, int32_t npairs,
                          const int32_t *const *inc_ptr,
                          const int32_t *const *inc_ent, int32_t set_diag,
                          int32_t *rowptr, int64_t *nnz_host, void *work,
                          size_t work_bytes, void *stream);
int op2hip_sparsity_fill(int32_t nrows, int32_t ncols,
                         const op2hip_map_pair *pairs, int32_t npairs,
                         const int32_t *const *inc_ptr,
                         const int32_t *const *inc_ent, int32_t set_diag,
                         const int32_t *rowptr, int32_t *colidx, void *stream);
/* Dof (scalar) CSR pattern of a block pattern with rdim x cdim blocks
 * (the AIJ layout of a non-block-sparse vector Mat, sparsity.pyx:168-173). */
int op2hip_pattern_expand(int32_t nrows, const int32_t *rowptr,
                          const int32_t *colidx, int32_t rdim, int32_t cdim,
                          int32_t *rowptr_out, int32_t *colidx_out,
                          void *stream);

/* ---- greedy colouring (doc/sphinx/source/plan.rst:46-80), host --------- */
/* Colours elements [0, nelem) so that no two elements of one colour share
 * an entry of any map (maps into the same target set share masks).  Host
 * arrays.  Returns the colour count in *ncolours. */
int op2hip_colour_greedy(int32_t nelem, int32_t nmaps,
                         const int32_t *const *maps, const int32_t *arity,
                         const int32_t *target_id, const int32_t *target_size,
                         int32_t ntargets, int32_t *colour, int32_t *ncolours);
/* Tiles of OP2HIP_STAGED (host arrays).  The elements of [start, end) are
 * grouped into tiles of at most tile_size elements and, when node_cap > 0,
 * at most node_cap distinct nodes of map 0; each tile grows breadth first
 * from the lowest unassigned element over elements sharing a node of map 0,
 * so tiles are compact regions.  Outputs (capacities for n = end-start):
 * *ntiles; tile_ptr (n+1: ntiles+1 used) and tile_elems (n element ids by
 * tile); per map m (nmaps <= 2, map entries >= 0): node_ptr[m] (n+1) and
 * nodes[m] (n*arity[m]): the distinct nodes of each tile, the nexcl[m][t]
 * (n) nodes no other tile references first, each part ascending; local[m]
 * (n*arity[m], tile_elems order): the position of map[e][i] among its
 * tile's nodes; max_nodes[m]: the largest node count of a tile.
 * tile_size * arity must stay below 65536. */
int op2hip_plan_tiles(int32_t start, int32_t end, int32_t nmaps, const int32_t *const *maps,
                      const int32_t *arity, const int32_t *target_size, int32_t tile_size,
                      int32_t node_cap, int32_t *ntiles, int32_t *tile_ptr,
                      int32_t *tile_elems, int32_t *const *node_ptr, int32_t *const *nodes,
                      int32_t *const *nexcl, uint16_t *const *local, int32_t *max_nodes);
/* Shared-node reduction plan of one staged map (op2hip_exec.tile_shared_off
 * and the arrays after it), from op2hip_plan_tiles' node_ptr / nodes / nexcl
 * of that map (host arrays).  The shared nodes of tile b (positions
 * nexcl[b].. of its node list) get partial slots shared_off[b] + k - nexcl[b];
 * shared_nodes / shared_ptr / shared_slots list, per distinct shared node in
 * increasing id order, its slots in tile order.  Capacities: shared_off
 * ntiles+1; shared_nodes, shared_slots: shared_off[ntiles] (the total of
 * node_ptr[b+1] - node_ptr[b] - nexcl[b]); shared_ptr one more.  No
 * counterpart in the reference, whose loop is sequential. */
int op2hip_plan_tile_reduce(int32_t ntiles, const int32_t *node_ptr, const int32_t *nodes,
                            const int32_t *nexcl, int32_t nnodes, int32_t *shared_off,
                            int32_t *nshared, int32_t *shared_nodes, int32_t *shared_ptr,
                            int32_t *shared_slots);
/* Row batches for OP2HIP_ROWGATHER: consecutive rows whose incidences and
 * blocks stay within max_inc / max_nnz (a single larger row gets its own
 * batch).  Host arrays; batch_rows has room for nrows+1 entries. */
int op2hip_plan_row_batches(int32_t nrows, const int32_t *inc_ptr,
                            const int32_t *rowptr, int32_t max_inc,
                            int32_t max_nnz, int32_t *batch_rows,
                            int32_t *nbatches, int32_t *max_batch_inc,
                            int32_t *max_batch_nnz, int32_t *max_batch_rows);

/* Slot table of OP2HIP_ROWGATHER_SLOTS for the batches of a row-batch plan:
 * per batch a descriptor {r0, nr, z0, nz, i0, ni, 0, 0} (8 int32) and, per
 * row incidence t (inc_ent order), an entry of entry_bytes(carity) bytes:
 * uint16 block offset of the row in its batch, uint16 row index in the
 * batch, then carity uint8 positions of column block j in the row (255: no
 * column or not in the pattern).  bad points to two device int32 counters,
 * zeroed first: [0] map entries not in the pattern, [1] row incidences whose
 * element names one column twice (OP2HIP_ROWGATHER_STAGED, which adds one
 * block per incidence and output, does not take such a table). */
int op2hip_slot_table_bytes(int32_t carity, size_t *entry_bytes);
int op2hip_build_slot_table(int32_t nbatches, const int32_t *batch_rows,
                            const int32_t *rowptr, const int32_t *colidx,
                            const int32_t *inc_ptr, const int32_t *inc_ent,
                            const int32_t *cmap, int32_t rarity, int32_t carity,
                            void *table, void *batch_desc, int32_t *bad,
                            void *stream);

/* Row-list plan of OP2HIP_ROWGATHER_ROWLIST (host arrays).  Rows are taken
 * in the order of their first incident element (ties by row id) and cut
 * into batches of at most max_inc incidences and max_nnz blocks (a longer
 * row is a batch of its own).  Outputs: row_order (2 * nrows int32: per
 * position, the row and its block offset within its batch); *nbatches;
 * desc (8 int32 per batch, capacity nrows: {first position in row_order,
 * rows, 0, blocks, first batch incidence, incidences, 0, 0}); per batch
 * incidence, in batch order (capacity the incidence count): ent (element *
 * rarity + i) and a slot entry of op2hip_slot_table_bytes(carity) bytes
 * (block offset of the row in the batch, row index in the batch, column
 * slots; 255: no column).  *missing counts map entries not in the pattern;
 * maxima of the batches go to max_inc_out / max_nnz_out / max_rows_out. */
int op2hip_plan_row_list(int32_t nrows, const int32_t *rowptr, const int32_t *colidx,
                         const int32_t *inc_ptr, const int32_t *inc_ent, const int32_t *cmap,
                         int32_t rarity, int32_t carity, int32_t max_inc, int32_t max_nnz,
                         int32_t *row_order, int32_t *nbatches, void *desc, int32_t *ent,
                         void *table, int64_t *missing, int32_t *max_inc_out,
                         int32_t *max_nnz_out, int32_t *max_rows_out);

/* Cell-gather plan of OP2HIP_ROWGATHER_CELLS (host arrays), from the batch
 * descriptors and slot table of op2hip_build_slot_table and the map
 * transpose's inc_ent, all copied to the host.  Per batch b: the distinct
 * elements of its row incidences, ascending, at cells[cell_ptr[b] ..]; per
 * Mat block z: gcount[z], the number of element blocks added to it; per
 * batch its blocks in thread order (order[z0 ..]: block index within the
 * batch, by count descending, ties ascending) and, in that order, each
 * block's element blocks at contrib[contrib_ptr[b] ..] as (u << 8) | (i << 4)
 * | j (u: position in the batch's cell list; i, j: local row and column), in
 * element order.  Capacities: cell_ptr and contrib_ptr nbatches+1, cells the
 * total incidence count, contrib that count times carity.  Batches of at
 * most 256 elements and 65536 blocks, arities <= 16, <= 255 element blocks
 * per Mat block (else OP2HIP_ERR_ARG). */
int op2hip_plan_cell_gather(int32_t nbatches, const void *batch_desc, const void *table,
                            int32_t rarity, int32_t carity, const int32_t *inc_ent,
                            int32_t *cell_ptr, int32_t *cells, uint8_t *gcount, uint16_t *order,
                            int32_t *contrib_ptr, uint16_t *contrib);

/* Gather table of OP2HIP_ROWGATHER_STAGED, from a slot table and its batch
 * descriptors (batches of at most 1024 blocks).  Output block z of a batch
 * receives the element blocks (t, j) -- row incidence t of the batch, column
 * block j -- whose slot is z.  Count: gcount[z0 + z] (uint8, one per Mat
 * block) = their number, batch_total[b] (device int32) = the batch's sum,
 * *over (device int32, zeroed first) = blocks with more than 255.  Fill,
 * given batch_off = exclusive prefix sum of batch_total: the batch's list at
 * glist + batch_off[b] holds t * carity + j (uint16) block by block, in
 * incidence order; the descriptors' last two fields become (batch_off[b],
 * batch_total[b]). */
int op2hip_gather_table_count(int32_t nbatches, int32_t batch_max_nnz, const void *batch_desc,
                              const void *table, int32_t carity, uint8_t *gcount,
                              int32_t *batch_total, int32_t *over, void *stream);
int op2hip_gather_table_fill(int32_t nbatches, int32_t batch_max_nnz, void *batch_desc,
                             const void *table, int32_t carity, const int32_t *batch_off,
                             const int32_t *batch_total, const uint8_t *gcount, uint16_t *glist,
                             void *stream);

/* ---- hand-written element loops ---------------------------------------- */
/* P1 triangles, op2.Kernel names p1_lumped_mass / p1_laplace_residual /
 * p1_laplace_matrix (pyop2_amd/elements.py).  X is (nverts, 2). */
int op2hip_wrap_p1_lumped_mass(int32_t start, int32_t end, double *M,
                               const double *X, const int32_t *c2v,
                               const op2hip_exec *ex, void *stream);
int op2hip_wrap_p1_laplace_residual(int32_t start, int32_t end, double *R,
                                    const double *X, const double *u,
                                    const int32_t *c2v, const op2hip_exec *ex,
                                    void *stream);
int op2hip_wrap_p1_laplace_matrix(int32_t start, int32_t end,
                                  const op2hip_mat *A, const double *X,
                                  const int32_t *c2v, const op2hip_exec *ex,
                                  void *stream);
/* P2 tetrahedra, isotropic elasticity (lam, mu), X (nverts, 3) through the
 * cell->vertex map c2v (arity 4), u and R (nnodes, 3) through the P2 map
 * c2n (arity 10).  Argument order follows the kernels' signatures:
 * residual (R, X, u; maps c2n, c2v), matrix (A, X; maps c2n, c2v), fused
 * (A, R, X, u; maps c2n, c2v).  _f32 variants take float data and a float
 * matrix. */
int op2hip_wrap_p2_elasticity_residual(int32_t start, int32_t end, double *R,
                                       const double *X, const double *u,
                                       const int32_t *c2n, const int32_t *c2v,
                                       double lam, double mu,
                                       const op2hip_exec *ex, void *stream);
int op2hip_wrap_p2_elasticity_residual_f32(int32_t start, int32_t end, float *R,
                                           const float *X, const float *u,
                                           const int32_t *c2n, const int32_t *c2v,
                                           double lam, double mu,
                                           const op2hip_exec *ex, void *stream);
int op2hip_wrap_p2_elasticity_matrix(int32_t start, int32_t end,
                                     const op2hip_mat *A, const double *X,
                                     const int32_t *c2n, const int32_t *c2v,
                                     double lam, double mu,
                                     const op2hip_exec *ex, void *stream);
int op2hip_wrap_p2_elasticity_matrix_f32(int32_t start, int32_t end,
                                         const op2hip_mat *A, const float *X,
                                         const int32_t *c2n, const int32_t *c2v,
                                         double lam, double mu,
                                         const op2hip_exec *ex, void *stream);
int op2hip_wrap_p2_elasticity_fused(int32_t start, int32_t end,
                                    const op2hip_mat *A, double *R,
                                    const double *X, const double *u,
                                    const int32_t *c2n, const int32_t *c2v,
                                    double lam, double mu,
                                    const op2hip_exec *ex, void *stream);
int op2hip_wrap_p2_elasticity_fused_f32(int32_t start, int32_t end,
                                        const op2hip_mat *A, float *R,
                                        const float *X, const float *u,
                                        const int32_t *c2n, const int32_t *c2v,
                                        double lam, double mu,
                                        const op2hip_exec *ex, void *stream);
/* The P2 residual loop without allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as op2hip_wrap_p2_elasticity_residual, plus
 * partial_slots, the partial-slot count of the tile plan's map 0 (the last
 * entry of tile_shared_off[0]).  Under OP2HIP_STAGED with a partial-sum plan
 * the partial sums go to a buffer inside the library (a device array of
 * OP2HIP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends) and
 * ex->partials are ignored; a plan needing more than that many bytes
 * (partial_slots * 3 doubles) is rejected with OP2HIP_ERR_ARG.  Other
 * strategies use no scratch.  Allocates nothing, synchronises nothing.
 * Every call on one GPU shares that GPU's buffer: calls that overlap on one
 * GPU (concurrent streams) race; nothing in it needs clearing between calls
 * (each slot is written by its tile before the second pass reads it). */
#define OP2HIP_NOALLOC_SCRATCH_BYTES (64u << 20)
int op2hip_wrap_p2_elasticity_residual_noalloc(int32_t start, int32_t end, double *R,
                                               const double *X, const double *u,
                                               const int32_t *c2n, const int32_t *c2v,
                                               double lam, double mu,
                                               const op2hip_exec *ex, int64_t partial_slots,
                                               void *stream);
/* Adjoint of the P2 residual loop: given Rbar (nnodes, 3), adds
 * ubar += K(X)^T Rbar and Xbar += d(Rbar . R(u, X))/dX (transposed
 * scatter/gather of the residual loop; no reference counterpart). */
int op2hip_wrap_p2_elasticity_adjoint(int32_t start, int32_t end,
                                      double *ubar, double *Xbar,
                                      const double *Rbar, const double *X,
                                      const double *u, const int32_t *c2n,
                                      const int32_t *c2v, double lam, double mu,
                                      const op2hip_exec *ex, void *stream);

/* ---- batched loops over many independent meshes (no reference counterpart:
 * the reference runs one par_loop per mesh) ---- */
/* One mesh of a batched P1 Laplace launch: CSR assembly of the Laplace
 * element matrices into A (slot-table row gather: A.inc_ptr/inc_ent from
 * the mesh's map transpose, batch descriptors and slot table from
 * op2hip_build_slot_table) and INC of the Laplace residual K u into R
 * (staged tiles of op2hip_plan_tiles over c2v).  All pointers device. */
typedef struct op2hip_batch_mesh {
  op2hip_mat A;
  const void *batch_desc;
  const void *slot_table;
  int32_t nbatches;
  int32_t ncells;
  const double *X;                /* vertex coordinates (nverts, 2) */
  const int32_t *c2v;             /* (ncells, 3) */
  const double *u;                /* (nverts) */
  double *R;                      /* (nverts), incremented */
  int32_t ntiles;
  int32_t max_tile_nodes;
  const int32_t *tile_ptr;
  const int32_t *tile_elems;
  const int32_t *tile_node_ptr;
  const int32_t *tile_nodes;
  const int32_t *tile_nexcl;
  const uint16_t *elem_local;
} op2hip_batch_mesh;

/* Work item flag: item (mesh, OP2HIP_BATCH_TILE | t) is residual tile t of
 * the mesh, item (mesh, b) row batch b of its assembly. */
#define OP2HIP_BATCH_TILE 0x40000000

/* Runs every work item (items: nitems (mesh, code) int32 pairs, device) of
 * the meshes (device array of op2hip_batch_mesh) in one launch of grid
 * workgroups (0: one per item) that stride over the items.  max_batch_nnz /
 * max_tile_nodes: largest batch / tile over the meshes (size the LDS).
 * Allocates nothing and synchronises nothing (safe under stream capture). */
int op2hip_batched_p1_laplace(const op2hip_batch_mesh *meshes, const int32_t *items,
                              int32_t nitems, int32_t grid, int32_t max_batch_nnz,
                              int32_t max_tile_nodes, void *stream);

/* ---- generic loops: runtime-compiled gather/compute/scatter template ---- */
/* Compiles HIP source (a generated wrapper around the user's C kernel) for
 * gfx950 with hiprtc and loads it.  *module receives an opaque handle. */
int op2hip_jit_compile(const char *source, const char *const *options,
                       int32_t noptions, void **module, char *log,
                       size_t log_bytes);
int op2hip_jit_code_object(const char *source, const char *const *options,
                           int32_t noptions, void *code, size_t *code_bytes,
                           char *log, size_t log_bytes);
int op2hip_jit_load(const void *code, void **module);
int op2hip_jit_unload(void *module);
/* Launches kernel `name` of `module`: args is an array of pointers to the
 * argument values (the hipModuleLaunchKernel convention). */
int op2hip_jit_launch(void *module, const char *name, uint32_t grid,
                      uint32_t block, uint32_t shared_bytes, void **args,
                      void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OP2HIP_H */
