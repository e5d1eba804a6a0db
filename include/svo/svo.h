/*
 * svo.h -- C ABI of the MI355X sparse-voxel-octree builder (libsvo_amd.so).
 *
 * A triangle mesh and a grid resolution go in; the octree node, neighbor,
 * vertex and edge arrays and the <stem>_<depth>_tree.obj file come out, with
 * the values the reference path computes (Alan-Leo-Wong/GPU-SVO):
 *
 *   reference entry point                         replaced by
 *   ----------------------------------------------------------------------------
 *   BaseModel::loadOBJ        BaseModel.cpp:3-25  svo_load_obj
 *   SparseVoxelOctree ctor    SVO.h:56-67         svo_mesh_desc + svo_params
 *   SparseVoxelOctree::create SVO.cu:341-520      svo_build / svo_build_device
 *     meshVoxelize            SVO.cu:190-206        (voxelize kernels)
 *     level loop              SVO.cu:383-495        (sort, unique, level kernels)
 *     constructNodeNeighbors  SVO.cu:548-565        (neighbor kernels)
 *     constructNodeVertexAndEdge SVO.cu:695-728     (vertex/edge kernels)
 *   private svoNodeArray      SVO.h:47            svo_tree_copy_nodes / _field
 *   private nodeVertexArray   SVO.h:50            svo_tree_copy_field(VERTICES)
 *   private nodeEdgeArray     SVO.h:51            svo_tree_copy_field(EDGES)
 *   private depthNumNodes     SVO.h:44            svo_tree_info.level_nodes
 *   SparseVoxelOctree::writeTree SVO.cu:737-753   svo_write_obj
 *   SparseVoxelOctree::writeVoxel SVO.cu:755-777  svo_write_voxels_obj
 *   main()                    main.cpp:7-17       the `svo` CLI (host/svo_main.cpp)
 *
 * Conventions: every function returns an svo_status (0 = success, negative =
 * error) and never terminates the process (the reference exits on errors,
 * CUDAUtil.h:21-48, BaseModel.cpp:4-8).  Pointers are plain host or device
 * pointers; sizes are in elements unless named *_bytes.  A context owns one
 * device and is safe to use from several threads; a tree is immutable after
 * its build and may be read from any thread.
 */
#ifndef SVO_SVO_H
#define SVO_SVO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SVO_OK = 0,
  SVO_E_INVALID = -1,   /* bad argument, bad face index, non-finite point, zero-extent bbox */
  SVO_E_RES_LIMIT = -2, /* resolution beyond what the key width can encode                  */
  SVO_E_OOM = -3,       /* device memory or context budget exhausted                         */
  SVO_E_HIP = -4,       /* HIP runtime error (no device, launch failure, ...)                */
  SVO_E_IO = -5,        /* file cannot be opened / written                                   */
  SVO_E_PARSE = -6,     /* OBJ line of 255+ characters (the reference loops forever on it)   */
  SVO_E_EMPTY = -7,     /* no points, no faces, or no voxel hit                              */
  SVO_E_BUSY = -8,      /* server queue full                                                 */
  SVO_E_RANGE = -9,     /* destination buffer too small                                      */
  SVO_E_TOO_LARGE = -10 /* input beyond the fixed scratch of a no-allocation entry point     */
} svo_status;

/* Output semantics.  REFERENCE reproduces the reference's arrays bit for bit,
 * including its neighbor table (only the root's centre slot is stored,
 * SVO.cu:528-542) and its vertex/edge lists (8 and 12 records per node,
 * SVO.cu:586-720).  INTENDED runs the top-down LUT neighbor lookup on the
 * stored array and keeps one record per distinct same-level vertex, edge and
 * face, owned by the sharing node with the smallest index. */
enum { SVO_SEM_REFERENCE = 0, SVO_SEM_INTENDED = 1 };

/* Per-node fields and output arrays (svo_tree_copy_field). */
typedef enum {
  SVO_FIELD_VOXELS = 0,      /* sorted unique level-1 Morton codes (uint32 or uint64)   */
  SVO_FIELD_CODE = 1,        /* node Morton code (uint32 or uint64)                       */
  SVO_FIELD_ISLEAF = 2,      /* uint8                                                     */
  SVO_FIELD_ORIGIN = 3,      /* float32 x3                                                */
  SVO_FIELD_WIDTH = 4,       /* float32                                                   */
  SVO_FIELD_PARENT = 5,      /* uint32, UINT32_MAX for the root                           */
  SVO_FIELD_CHILDS = 6,      /* uint32 x8                                                 */
  SVO_FIELD_NEIGHBORS = 7,   /* uint32 x27, slot = (dx+1) + 3(dy+1) + 9(dz+1)             */
  SVO_FIELD_VERTICES = 8,    /* svo_vertex records                                        */
  SVO_FIELD_EDGES = 9,       /* svo_edge records                                          */
  SVO_FIELD_FACES = 10,      /* svo_edge records {lo corner, hi corner, owner} (INTENDED) */
  SVO_FIELD_LEVEL_NODES = 11,/* uint64 nodes per level, finest level first                */
  SVO_FIELD_NODES_AOS = 12,  /* svo_node_rec, the reference's 168-byte SVONode layout     */
  SVO_FIELD_COUNT = 13
} svo_field;

typedef struct {
  float p[3];
  uint32_t owner;
} svo_vertex; /* thrust::pair<Eigen::Vector3f, uint32_t>, SVO.h:50 (16 bytes) */

typedef struct {
  float a[3];
  float b[3];
  uint32_t owner;
} svo_edge; /* thrust::pair<thrust_edge, uint32_t>, SVO.h:37,51 (28 bytes) */

typedef struct {
  uint32_t morton_code;
  uint8_t is_leaf;
  uint8_t pad[3]; /* written as zero */
  float origin[3];
  float width;
  uint32_t parent;
  uint32_t childs[8];
  uint32_t neighbors[27];
} svo_node_rec; /* SVONode, SVO.h:20-35 (168 bytes) */

typedef struct svo_ctx svo_ctx;
typedef struct svo_mesh svo_mesh;
typedef struct svo_tree svo_tree;
typedef struct svo_server svo_server;

typedef struct {
  const float* xyz;   /* n_points x 3 float32 */
  uint64_t n_points;
  const int32_t* tri; /* n_tris x 3 zero-based vertex indices */
  uint64_t n_tris;
} svo_mesh_desc;

typedef struct {
  int32_t res[3];     /* surfaceVoxelGridSize, SVO.h:43 */
  uint32_t semantics; /* SVO_SEM_* */
  uint32_t key_bits;  /* 32 (reference, res <= 1024), 64 (res <= 65536), 0 = pick */
  uint32_t flags;     /* SVO_BUILD_* */
} svo_params;

enum {
  SVO_BUILD_NO_VERTEX_EDGE = 1u << 0, /* skip vertex/edge/face arrays */
};

typedef struct {
  int32_t depth;            /* treeDepth, SVO.h:41 */
  int32_t key_bits;
  int32_t semantics;
  int32_t res[3];
  float bbox_min[3];        /* modelBBox after createMeshBBCube, BaseModel.h:24-49 */
  float bbox_max[3];
  float bbox_width[3];
  float unit[3];            /* unitVoxelSize, SVO.cu:348-350 */
  uint64_t n_voxels;
  uint64_t n_nodes;         /* numTreeNodes */
  uint64_t n_vertices;
  uint64_t n_edges;
  uint64_t n_faces;
  uint64_t n_hits;          /* voxel hits before de-duplication */
  uint64_t device_bytes;    /* device memory held by the tree's arrays */
  uint64_t level_nodes[32]; /* nodes per level, finest first (depthNumNodes + root) */
} svo_tree_info;

/* ---- context ------------------------------------------------------------ */
/* mem_budget: upper bound on device bytes the context may allocate (0 = no
 * bound beyond the device). */
int svo_ctx_create(int device, uint64_t mem_budget, svo_ctx** out);
int svo_ctx_destroy(svo_ctx* ctx);
/* Per-stage device times of the last build call on this context, in
 * milliseconds, measured with HIP events on the builds' streams when timing
 * is enabled.  A call that runs as several concurrent passes reports, per
 * stage, the window from the stage's earliest start to its latest end. */
int svo_ctx_set_timing(svo_ctx* ctx, int enable);
int svo_ctx_last_stage_ms(svo_ctx* ctx, double* ms, const char** names, int cap, int* n);
/* Batch plans (no counterpart in the reference; they replace the per-level
 * host synchronizations of its level loop, SVO.cu:383-495, by one per
 * batch).  svo_plan_create builds the batch once eagerly, recording the
 * sizes every pass reads back, then captures the whole batched build -- the
 * sub-batch passes as parallel branches -- in one HIP graph.  Host meshes are
 * copied to device memory the plan owns; device meshes must stay allocated
 * while the plan lives.  All meshes of a plan share one semantics and flag
 * set.  svo_plan_run replays the graph (one launch, one synchronization),
 * recomputing every output from the current mesh contents; when a count,
 * span or total differs from the recorded build (the mesh contents changed
 * their sizes) it builds eagerly instead and sets *replayed to 0.  The trees
 * of a replay share the plan's output blocks: they stay valid until the
 * next svo_plan_run or svo_plan_destroy (free them with svo_tree_free).
 * Calls on one plan must not overlap. */
typedef struct svo_plan svo_plan;
int svo_plan_create(svo_ctx* ctx, const svo_mesh_desc* meshes, const svo_params* params,
                    uint32_t n, int meshes_on_device, svo_plan** out);
int svo_plan_run(svo_plan* plan, svo_tree** trees, int* statuses, int* replayed);
int svo_plan_destroy(svo_plan* plan);

/* Tuning options of the context (no counterpart in the reference), set
 * between builds:
 *   "small_max_cells"        box cells up to which one lane voxelizes a triangle (1024)
 *   "sort_blocks"            grid bound of the radix sort passes (65536)
 *   "scan_blocks"            persistent grid bound of the level passes (1024)
 *   "bitmap_max_bytes"       largest occupancy bitmap used instead of the sort (32 GiB)
 *   "bitmap_bytes_per_cell"  bitmap bytes allowed per candidate cell (64)
 *   "dedup_path"             0 automatic, 1 always the bitmap, 2 always the sort
 *   "batch_lanes"            concurrent passes a large svo_build_batch is split into (2)
 *   "batch_min_meshes"       fewest meshes per such pass (16)
 * Unknown names or out-of-range values give SVO_E_INVALID. */
int svo_ctx_set_option(svo_ctx* ctx, const char* name, int64_t value);
int svo_ctx_get_option(svo_ctx* ctx, const char* name, int64_t* value);
/* Tuning search (no counterpart in the reference), on a sample batch:
 *  1. builds it `reps` times (after one warm-up; the median counts) under the
 *     current options and under each combination of small_max_cells in
 *     {128, 256, 512, 1024, 2048} and batch_lanes in {1, 2, 3, 4}, and keeps
 *     the fastest;
 *  2. then, per size class present in the sample (5 resolution bands x 3
 *     bands of res^2 / triangles; the option names class_small_max_<k>),
 *     builds that class's meshes alone with the one-lane box limit of the
 *     class in {64, 128, 256, 512, 1024, 2048, 4096} and keeps the limit
 *     whose voxelization (classify + voxelize stages, HIP events) is
 *     fastest, when it beats the global limit by more than 3 %.
 * The choices stay on the context; the trees built are freed. */
typedef struct {
  uint32_t small_max_cells, batch_lanes; /* chosen */
  double best_ms, default_ms;            /* median wall ms per build of the sample */
  uint32_t n_trials;
  uint32_t class_small_max[15];          /* per size class; 0 = small_max_cells */
  uint32_t class_meshes[15];             /* sample meshes of each class */
} svo_tune_result;
int svo_ctx_autotune(svo_ctx* ctx, const svo_mesh_desc* meshes, const svo_params* params,
                     uint32_t n, int meshes_on_device, uint32_t reps, svo_tune_result* out);
/* Options as "name value" lines (every option svo_ctx_set_option knows). */
int svo_ctx_save_tuning(svo_ctx* ctx, const char* path);
int svo_ctx_load_tuning(svo_ctx* ctx, const char* path);
/* Sizes of the last build call on this context, summed over its passes. */
typedef struct {
  uint64_t n_meshes, n_tris, box_cells, n_hits, n_voxels, level_elems, n_nodes;
  uint64_t n_vertices, n_edges, n_faces, out_bytes;
  uint32_t key_bytes, sort_bits, sort_passes, n_levels;
  uint64_t bitmap_bytes; /* occupancy bitmap of the pass; 0 when the hits were sorted */
  uint64_t n_groups1;    /* parent groups of the unique voxels (level-2 elements) */
  uint64_t bitmap_fill_bytes;  /* bitmap bytes cleared before voxelization (0 when the
                                  lane's bitmap was already clear) */
  uint64_t bitmap_reset_bytes; /* bitmap bytes zeroed by the level-1 pass after reading */
} svo_pass_stats;
int svo_ctx_last_pass_stats(svo_ctx* ctx, svo_pass_stats* out);
uint64_t svo_ctx_bytes_in_use(const svo_ctx* ctx);
/* The part of svo_ctx_bytes_in_use held by the pipeline lanes' workspaces,
 * which are kept between builds; the rest is trees and svo_device_alloc. */
uint64_t svo_ctx_workspace_bytes(const svo_ctx* ctx);
/* Device allocations (hipMalloc calls) the context has made: blocks freed by
 * the library are cached and reused, so in a steady loop of builds of one
 * shape this count stops growing after the first builds. */
uint64_t svo_ctx_device_mallocs(const svo_ctx* ctx);
/* Device memory on the context's device (counted against its budget), for
 * callers that keep meshes resident for svo_build_device / svo_build_batch. */
int svo_device_alloc(svo_ctx* ctx, uint64_t bytes, void** out);
int svo_device_free(svo_ctx* ctx, void* ptr, uint64_t bytes);
/* kind: 0 = host to device, 1 = device to host, 2 = device to device. */
int svo_device_copy(svo_ctx* ctx, void* dst, const void* src, uint64_t bytes, int kind);

/* ---- meshes ----------------------------------------------------------- */
/* BaseModel::loadOBJ rules (BaseModel.cpp:3-25): "v x y z" lines; faces
 * "f a//b c//d e//f", "f a/b c/d e/f" or "f a b c" (first three indices);
 * other face forms are dropped; a last line without newline is dropped. */
int svo_load_obj(const char* path, svo_mesh** out);
int svo_mesh_view(const svo_mesh* mesh, svo_mesh_desc* out);
void svo_mesh_free(svo_mesh* mesh);
/* Checks a build request on the host without touching the device. */
int svo_validate(const svo_mesh_desc* mesh, const svo_params* params);

/* ---- builds ----------------------------------------------------------- */
/* Mesh arrays in host memory. */
int svo_build(svo_ctx* ctx, const svo_mesh_desc* mesh, const svo_params* params, svo_tree** out);
/* Mesh arrays already in device memory of the context's device. */
int svo_build_device(svo_ctx* ctx, const svo_mesh_desc* mesh, const svo_params* params,
                     svo_tree** out);
/* Many meshes in one pipeline pass (composite (mesh, Morton) keys); meshes
 * with different semantics or flags run as one pass per (semantics, flags).
 * meshes_on_device: 0 = host arrays, 1 = device arrays.  statuses may be NULL.
 * Returns SVO_OK if every mesh built; per-mesh results in statuses[i]. */
int svo_build_batch(svo_ctx* ctx, const svo_mesh_desc* meshes, const svo_params* params,
                    uint32_t n, int meshes_on_device, svo_tree** trees, int* statuses);

/* ---- trees ------------------------------------------------------------ */
int svo_tree_info_get(const svo_tree* tree, svo_tree_info* out);
uint64_t svo_tree_field_bytes(const svo_tree* tree, int field);
/* Copies a field to host memory. */
int svo_tree_copy_field(const svo_tree* tree, int field, void* dst, uint64_t dst_bytes);
/* Named read-outs of the reference's arrays (each svo_tree_copy_field of one
 * field; dst_bytes at least svo_tree_field_bytes of it):
 *   voxels     sorted unique Morton codes (compactArray output, SVO.cu:405)
 *   nodes      svo_node_rec[n_nodes] (svoNodeArray, SVO.h:47)
 *   neighbors  uint32[n_nodes][27] (SVONode::neighbors, SVO.h:33)
 *   vertices   svo_vertex[n_vertices] (nodeVertexArray, SVO.h:50)
 *   edges      svo_edge[n_edges] (nodeEdgeArray, SVO.h:51)
 *   faces      svo_edge[n_faces] (intended semantics only) */
int svo_tree_copy_voxels(const svo_tree* tree, void* dst, uint64_t dst_bytes);
int svo_tree_copy_nodes(const svo_tree* tree, svo_node_rec* dst, uint64_t dst_bytes);
int svo_tree_copy_neighbors(const svo_tree* tree, uint32_t* dst, uint64_t dst_bytes);
int svo_tree_copy_vertices(const svo_tree* tree, svo_vertex* dst, uint64_t dst_bytes);
int svo_tree_copy_edges(const svo_tree* tree, svo_edge* dst, uint64_t dst_bytes);
int svo_tree_copy_faces(const svo_tree* tree, svo_edge* dst, uint64_t dst_bytes);
/* Device pointer of a field, valid until svo_tree_free (NULL for AOS/LEVEL). */
const void* svo_tree_device_ptr(const svo_tree* tree, int field);
/* writeTree, SVO.cu:737-753 + IO.h:23-90: writes <dir>/<stem>_<depth>_tree.obj
 * (dir may be NULL or "" for the working directory); the path written is
 * returned in path_out (may be NULL). */
int svo_write_obj(const svo_tree* tree, const char* dir, const char* stem, char* path_out,
                  size_t path_cap);
/* writeVoxel, SVO.cu:755-777 (which create() calls in NDEBUG builds,
 * SVO.cu:411-419, always with the stem "bunny"): writes
 * <dir>/<stem>_1_voxel.obj, one cube per occupied voxel in Morton order at
 * bbox_min + unit[0] * (x, y, z) with edge unit[0], in the writeTree format. */
int svo_write_voxels_obj(const svo_tree* tree, const char* dir, const char* stem, char* path_out,
                         size_t path_cap);
/* Neighbor lookup on a resident tree: out[i] = neighbors[node[i]][slot[i]].
 * node, slot and out are host arrays. */
int svo_query_neighbors(const svo_tree* tree, const uint32_t* node, const uint8_t* slot,
                        uint64_t n, uint32_t* out);
void svo_tree_free(svo_tree* tree);

const char* svo_status_str(int status);

/* ---- served path: continuous batching under a device-memory budget ------ */
typedef struct {
  uint64_t mem_budget;   /* device bytes for resident trees + in-flight work */
  uint32_t max_batch;    /* meshes per pipeline pass                         */
  uint32_t max_queue;    /* pending requests before submit returns SVO_E_BUSY */
  uint32_t n_lanes;      /* concurrent pipeline passes (one stream each)      */
  uint32_t max_wait_us;  /* batching window after the first pending request   */
} svo_server_cfg;

typedef struct {
  uint64_t request_id;
  int32_t status;
  uint32_t batch_size;   /* meshes in the pass that built this request        */
  svo_tree* tree;        /* owned by the caller after poll (svo_tree_free)    */
  double t_submit_s, t_start_s, t_done_s; /* steady clock, seconds           */
} svo_completion;

int svo_server_create(svo_ctx* ctx, const svo_server_cfg* cfg, svo_server** out);
/* Copies the mesh arrays (host memory); returns the request id. */
int svo_server_submit(svo_server* srv, const svo_mesh_desc* mesh, const svo_params* params,
                      uint64_t* request_id);
/* Returns up to cap completions; waits up to timeout_us for the first one. */
int svo_server_poll(svo_server* srv, svo_completion* out, uint32_t cap, uint32_t* n,
                    int64_t timeout_us);
/* Quick device-memory estimate from the counts alone, for a closed surface
 * spanning its box (not a bound: a triangle soup can take far more). */
uint64_t svo_estimate_bytes(uint64_t n_points, uint64_t n_tris, const svo_params* params);
/* Upper bound on the device bytes one build of the mesh takes on a fresh
 * context (output plus every workspace buffer, allocator rounding included),
 * from one host pass over its points and triangles; the served path admits
 * requests by it.  Inputs the build rejects get svo_estimate_bytes. */
uint64_t svo_estimate_bytes_mesh(const svo_mesh_desc* mesh, const svo_params* params);
int svo_server_stats(svo_server* srv, uint64_t* bytes_reserved, uint32_t* pending,
                     uint32_t* in_flight);
int svo_server_destroy(svo_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* SVO_SVO_H */
