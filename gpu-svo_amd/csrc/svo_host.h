// svo_host.h -- host-side objects behind the C ABI: context, pipeline lanes,
// workspaces, batch outputs and trees.
#pragma once
#include <atomic>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "svo_internal.h"

namespace svo {

struct Context;

// Named device buffers of one pipeline lane, kept between passes: a buffer
// grows when a pass needs more and is given back when it is far larger than
// a pass needs.
struct Workspace {
  struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
    size_t clean = 0;  // leading bytes known to be zero (kept by the caller)
  };
  Context* ctx = nullptr;
  std::unordered_map<std::string, Buf> bufs;
  // A frozen workspace hands out its buffers as they are and allocates
  // nothing (a pass captured into a graph): a buffer missing or too small is
  // an error.
  bool frozen = false;
  int get(const char* name, size_t bytes, void** out);
  // Leading bytes of buffer `name` known to be zero; 0 after (re)allocation.
  size_t& clean_bytes(const char* name) { return bufs[name].clean; }
  void release_all();
  size_t total() const;
};

// One stream and its workspace; one build runs on a lane at a time.  `side`
// runs work of the pass that overlaps the main stream (forked and joined
// with `fork` / `join`).
struct Lane {
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  Workspace ws;
  void* pinned = nullptr;  // host readback area
  size_t pinned_bytes = 0;
  std::vector<hipEvent_t> events;
  int ensure_pinned(size_t bytes);
};

// Refcounted device allocation holding a batch's output arrays.  It keeps
// its context alive: a tree may outlive svo_ctx_destroy.
struct DeviceBlock {
  std::shared_ptr<Context> ctx;
  void* ptr = nullptr;
  size_t bytes = 0;
  ~DeviceBlock();
};

struct BatchOutput {
  std::shared_ptr<DeviceBlock> block;
  NodeArrays nodes{};
  svo_vertex* verts = nullptr;
  svo_edge* edges = nullptr;
  svo_edge* faces = nullptr;
  void* voxels = nullptr;
  int code_bytes = 4;
};

// The device state behind an svo_ctx.  Shared: the caller's handle, every
// tree (through its output block) and every server hold a reference, so
// trees stay readable after svo_ctx_destroy, which only gives the lanes
// (streams and workspaces) back; the rest goes with the last reference.
struct Context : std::enable_shared_from_this<Context> {
  int device = 0;
  uint64_t budget = 0;  // 0 = unbounded
  std::atomic<uint64_t> in_use{0};
  std::atomic<uint64_t> ws_bytes{0};  // part of in_use held by lane workspaces (reused)
  std::atomic<int> active_passes{0};  // pipeline passes running on the context
  std::atomic<int> servers{0};        // live svo_server objects on the context
  Tuning tuning;                      // guarded by mu; a pass works on a snapshot
  std::mutex mu;
  std::condition_variable cv;
  std::vector<std::unique_ptr<Lane>> lanes;
  std::vector<Lane*> free_lanes;
  hipStream_t aux_stream = nullptr;  // copies for read-out
  std::mutex aux_mu;
  std::atomic<bool> timing{false};
  std::mutex stage_mu;
  std::vector<std::string> stage_names;
  std::vector<double> stage_ms;
  svo_pass_stats last_pass{};

  int init(int dev, uint64_t budget);
  // Gives the lanes back (synchronizes and destroys their streams, frees
  // their workspaces) and trims the block cache; idempotent.
  void release_lanes();
  ~Context();
  static constexpr size_t MAX_LANES = 8;
  // A free lane, lane `prefer` if it is free (made on demand, at most
  // MAX_LANES); waits while all lanes are busy.
  Lane* acquire_lane(int prefer = -1);
  void release_lane(Lane* l);
  int reserve(uint64_t bytes);  // budget accounting; SVO_E_OOM if over
  void release(uint64_t bytes);
  // Device allocations go through a size-class cache: freed blocks are kept
  // for reuse (hipFree synchronizes the device) up to cache_limit bytes.
  int dev_malloc(void** p, size_t bytes);
  void dev_free(void* p, size_t bytes);
  void trim_cache(uint64_t keep_bytes);
  static size_t size_class(size_t bytes);
  std::mutex cache_mu;
  std::multimap<size_t, void*> cache;
  uint64_t cached_bytes = 0;
  std::atomic<uint64_t> n_mallocs{0};  // hipMalloc calls (cache misses)
  uint64_t cache_limit = 96ull << 30;
};

}  // namespace svo

struct svo_ctx {
  std::shared_ptr<svo::Context> impl;
};

struct svo_mesh {
  std::vector<float> points;
  std::vector<int32_t> tris;
};

struct svo_tree {
  std::shared_ptr<svo::BatchOutput> out;
  std::shared_ptr<svo::Context> ctx;
  svo_tree_info info{};
  uint64_t node_base = 0;
  uint64_t vox_base = 0;
  uint64_t v_base = 0, e_base = 0, f_base = 0;
  int key_bits_out = 32;
  int semantics = 0;
};

namespace svo {
// The host values one pass reads back at its three synchronization points,
// kept by a batch plan (api: svo_plan_*): recorded from an eager run, they
// let the same pass be captured into a HIP graph with no readback at all;
// the captured pass compares what it computes with them on the device and
// sets a flag when anything differs.
struct PassRecord {
  enum Mode { RECORD, CAPTURE } mode = RECORD;
  uint64_t a0[2] = {0, 0};  // candidate cells, big-triangle work items
  std::vector<char> a1;     // level readback: counts, spans, states, params, hits
  std::vector<uint64_t> tot;  // record totals and per-mesh record offsets
  std::shared_ptr<BatchOutput> out;  // output block, written by every replay
  // CAPTURE: expected readbacks on the device ([a0 | a1 | tot]) and the flag
  const char* d_expect = nullptr;
  uint32_t* d_flag = nullptr;
  // CAPTURE: the trees a replay hands out (copied per replay), by input index
  std::vector<svo_tree> templ;
  std::vector<int> templ_status;
};
}  // namespace svo

namespace svo {

struct MeshInput {
  svo_mesh_desc desc;
  svo_params params;
  bool on_device;
};

// Runs the pipeline over the meshes.  trees[i] receives a tree or nullptr;
// statuses[i] the per-mesh status.  Returns SVO_OK unless the pass itself
// failed (HIP error or out of memory).
int run_batch(Context& C, const std::vector<MeshInput>& in, std::vector<svo_tree*>& trees,
              std::vector<int>& statuses);

// One pipeline pass: which meshes, their key layout and the tuning snapshot.
struct CallTiming;
struct Plan {
  CallTiming* timing = nullptr;
  Tuning tuning;              // snapshot taken under the context's mutex
  std::vector<uint32_t> idx;  // active mesh -> input index
  int bits = 1;               // per-axis coordinate bits of the batch
  int cb = 3;                 // code bits
  int mb = 0;                 // mesh bits
  bool k64 = false;
  int n_levels = 2;
  int code_bytes = 4;
  PassRecord* rec = nullptr;  // set by a batch plan (record or capture)
};

// A recorded batch replayed as one HIP graph (pipeline.cpp).
struct BatchPlan {
  std::shared_ptr<Context> ctx;
  std::vector<MeshInput> in;  // device-resident inputs
  std::vector<std::pair<void*, size_t>> owned;   // device copies of host inputs
  std::vector<std::unique_ptr<Lane>> lanes;      // one per sub-batch pass
  std::vector<std::vector<uint32_t>> sel;        // meshes of each pass
  std::vector<Plan> plans;
  std::vector<std::unique_ptr<PassRecord>> recs;
  std::vector<std::pair<void*, size_t>> expect;  // expected readbacks on the device
  uint32_t* d_flag = nullptr;
  uint32_t* h_flag = nullptr;  // pinned
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::mutex mu;
  ~BatchPlan();
};
int plan_create(Context& C, const std::vector<MeshInput>& in, std::unique_ptr<BatchPlan>& out);
int plan_run(BatchPlan& bp, std::vector<svo_tree*>& trees, std::vector<int>& statuses,
             bool& replayed);

int validate_params(const svo_mesh_desc* mesh, const svo_params* p, int* key_bits_out);
uint64_t estimate_bytes(uint64_t n_points, uint64_t n_tris, const svo_params* p);
uint64_t estimate_bytes_mesh(const svo_mesh_desc* m, const svo_params* p);
uint64_t grid_bound_bytes(const svo_params* p, int key_bits, uint64_t n_points, uint64_t n_tris);

// OBJ I/O (obj_io.cpp).
int load_obj_file(const char* path, std::vector<float>& pts, std::vector<int32_t>& tris);
int write_tree_obj(const char* path, const float* origin, const float* width, uint64_t n);

}  // namespace svo
