// api.cpp -- the C ABI declared in include/svo/svo.h.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "svo_host.h"

using namespace svo;

namespace {

struct FieldRange {
  const void* dev = nullptr;
  uint64_t elem_bytes = 0;
  uint64_t count = 0;
  bool is_code = false;
};

FieldRange field_range(const svo_tree* t, int field) {
  FieldRange r;
  const BatchOutput& o = *t->out;
  const NodeArrays& n = o.nodes;
  const uint64_t N = t->info.n_nodes, b = t->node_base;
  const uint64_t cb = (uint64_t)o.code_bytes;
  switch (field) {
    case SVO_FIELD_VOXELS:
      r.dev = (const char*)o.voxels + cb * t->vox_base;
      r.elem_bytes = cb;
      r.count = t->info.n_voxels;
      r.is_code = true;
      break;
    case SVO_FIELD_CODE:
      r.dev = (const char*)n.code + cb * b;
      r.elem_bytes = cb;
      r.count = N;
      r.is_code = true;
      break;
    case SVO_FIELD_ISLEAF: r.dev = n.is_leaf + b; r.elem_bytes = 1; r.count = N; break;
    case SVO_FIELD_ORIGIN: r.dev = n.origin + 3 * b; r.elem_bytes = 12; r.count = N; break;
    case SVO_FIELD_WIDTH: r.dev = n.width + b; r.elem_bytes = 4; r.count = N; break;
    case SVO_FIELD_PARENT: r.dev = n.parent + b; r.elem_bytes = 4; r.count = N; break;
    case SVO_FIELD_CHILDS: r.dev = n.childs + 8 * b; r.elem_bytes = 32; r.count = N; break;
    case SVO_FIELD_NEIGHBORS: r.dev = n.neighbors + 27 * b; r.elem_bytes = 108; r.count = N; break;
    case SVO_FIELD_VERTICES:
      r.dev = o.verts + t->v_base; r.elem_bytes = 16; r.count = t->info.n_vertices; break;
    case SVO_FIELD_EDGES:
      r.dev = o.edges + t->e_base; r.elem_bytes = 28; r.count = t->info.n_edges; break;
    case SVO_FIELD_FACES:
      r.dev = o.faces + t->f_base; r.elem_bytes = 28; r.count = t->info.n_faces; break;
    default: break;
  }
  return r;
}

int copy_d2h(Context& C, void* dst, const void* src, size_t bytes) {
  if (!bytes) return SVO_OK;
  std::lock_guard<std::mutex> lk(C.aux_mu);
  hipSetDevice(C.device);
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C.aux_stream) != hipSuccess)
    return SVO_E_HIP;
  if (hipStreamSynchronize(C.aux_stream) != hipSuccess) return SVO_E_HIP;
  return SVO_OK;
}

}  // namespace

extern "C" {

const char* svo_status_str(int s) {
  switch (s) {
    case SVO_OK: return "ok";
    case SVO_E_INVALID: return "invalid argument or input";
    case SVO_E_RES_LIMIT: return "resolution beyond the key width";
    case SVO_E_OOM: return "out of device memory or budget";
    case SVO_E_HIP: return "HIP runtime error";
    case SVO_E_IO: return "file cannot be opened or written";
    case SVO_E_PARSE: return "OBJ line of 255 or more characters";
    case SVO_E_EMPTY: return "empty mesh or no voxel";
    case SVO_E_BUSY: return "server queue full";
    case SVO_E_RANGE: return "destination buffer too small";
    case SVO_E_TOO_LARGE: return "input beyond the fixed scratch of a no-allocation entry point";
    default: return "unknown status";
  }
}

int svo_ctx_create(int device, uint64_t mem_budget, svo_ctx** out) {
  if (!out) return SVO_E_INVALID;
  *out = nullptr;
  svo_ctx* c = new svo_ctx();
  c->impl = std::make_shared<Context>();
  const int st = c->impl->init(device, mem_budget);
  if (st != SVO_OK) {
    delete c;
    return st;
  }
  *out = c;
  return SVO_OK;
}

int svo_ctx_destroy(svo_ctx* ctx) {
  if (!ctx) return SVO_E_INVALID;
  // servers drive the lanes from their own threads: destroy them first
  if (ctx->impl->servers.load() > 0) return SVO_E_BUSY;
  ctx->impl->release_lanes();
  delete ctx;  // trees still alive keep the rest of the device state
  return SVO_OK;
}

int svo_ctx_set_timing(svo_ctx* ctx, int enable) {
  if (!ctx) return SVO_E_INVALID;
  ctx->impl->timing = enable != 0;
  return SVO_OK;
}

namespace {
// name -> (field, min, max)
struct OptionRef {
  const char* name;
  int64_t lo, hi;
};
const OptionRef kOptions[] = {
    {"small_max_cells", 1, 1 << 20},
    {"sort_blocks", 1, 1 << 16},
    {"scan_blocks", 1, 1 << 16},
    {"bitmap_max_bytes", 0, INT64_MAX},
    {"bitmap_bytes_per_cell", 0, 1 << 20},
    {"dedup_path", 0, 2},
    {"batch_lanes", 1, 8},
    {"batch_min_meshes", 1, 1 << 20},
    {"class_small_max_0", 0, 1 << 20},  {"class_small_max_1", 0, 1 << 20},
    {"class_small_max_2", 0, 1 << 20},  {"class_small_max_3", 0, 1 << 20},
    {"class_small_max_4", 0, 1 << 20},  {"class_small_max_5", 0, 1 << 20},
    {"class_small_max_6", 0, 1 << 20},  {"class_small_max_7", 0, 1 << 20},
    {"class_small_max_8", 0, 1 << 20},  {"class_small_max_9", 0, 1 << 20},
    {"class_small_max_10", 0, 1 << 20}, {"class_small_max_11", 0, 1 << 20},
    {"class_small_max_12", 0, 1 << 20}, {"class_small_max_13", 0, 1 << 20},
    {"class_small_max_14", 0, 1 << 20},
};
static_assert(SIZE_CLASSES == 15, "one class_small_max option per size class");

// class index of a "class_small_max_<k>" name, or -1
int class_option(const std::string& n) {
  const std::string p = "class_small_max_";
  if (n.compare(0, p.size(), p) != 0) return -1;
  const int k = std::atoi(n.c_str() + p.size());
  return (k >= 0 && k < SIZE_CLASSES) ? k : -1;
}
}  // namespace

int svo_ctx_set_option(svo_ctx* ctx, const char* name, int64_t value) {
  if (!ctx || !name) return SVO_E_INVALID;
  Context& C = *ctx->impl;
  std::lock_guard<std::mutex> lk(C.mu);
  Tuning& t = C.tuning;
  for (const OptionRef& o : kOptions) {
    if (std::strcmp(o.name, name) != 0) continue;
    if (value < o.lo || value > o.hi) return SVO_E_INVALID;
    const std::string n(name);
    if (n == "small_max_cells") t.small_max_cells = (uint32_t)value;
    else if (n == "sort_blocks") t.sort_blocks = (uint32_t)value;
    else if (n == "scan_blocks") t.scan_blocks = (uint32_t)value;
    else if (n == "bitmap_max_bytes") t.bitmap_max_bytes = (uint64_t)value;
    else if (n == "bitmap_bytes_per_cell") t.bitmap_bytes_per_cell = (uint32_t)value;
    else if (n == "batch_lanes") t.batch_lanes = (uint32_t)value;
    else if (n == "batch_min_meshes") t.batch_min_meshes = (uint32_t)value;
    else if (class_option(n) >= 0) t.class_small_max[class_option(n)] = (uint32_t)value;
    else t.force_path = (int)value;
    return SVO_OK;
  }
  return SVO_E_INVALID;
}

int svo_ctx_get_option(svo_ctx* ctx, const char* name, int64_t* value) {
  if (!ctx || !name || !value) return SVO_E_INVALID;
  Context& C = *ctx->impl;
  std::lock_guard<std::mutex> lk(C.mu);
  const Tuning& t = C.tuning;
  const std::string n(name);
  if (n == "small_max_cells") *value = t.small_max_cells;
  else if (n == "sort_blocks") *value = t.sort_blocks;
  else if (n == "scan_blocks") *value = t.scan_blocks;
  else if (n == "bitmap_max_bytes") *value = (int64_t)t.bitmap_max_bytes;
  else if (n == "bitmap_bytes_per_cell") *value = t.bitmap_bytes_per_cell;
  else if (n == "dedup_path") *value = t.force_path;
  else if (n == "batch_lanes") *value = t.batch_lanes;
  else if (n == "batch_min_meshes") *value = t.batch_min_meshes;
  else if (class_option(n) >= 0) *value = t.class_small_max[class_option(n)];
  else return SVO_E_INVALID;
  return SVO_OK;
}

int svo_ctx_last_stage_ms(svo_ctx* ctx, double* ms, const char** names, int cap, int* n) {
  if (!ctx || !n) return SVO_E_INVALID;
  Context& C = *ctx->impl;
  std::lock_guard<std::mutex> lk(C.stage_mu);
  const int k = (int)C.stage_ms.size();
  *n = k;
  for (int i = 0; i < k && i < cap; ++i) {
    if (ms) ms[i] = C.stage_ms[i];
    if (names) names[i] = C.stage_names[i].c_str();
  }
  return k > cap ? SVO_E_RANGE : SVO_OK;
}

uint64_t svo_ctx_bytes_in_use(const svo_ctx* ctx) { return ctx ? ctx->impl->in_use.load() : 0; }

uint64_t svo_ctx_workspace_bytes(const svo_ctx* ctx) {
  return ctx ? ctx->impl->ws_bytes.load() : 0;
}

uint64_t svo_ctx_device_mallocs(const svo_ctx* ctx) {
  return ctx ? ctx->impl->n_mallocs.load() : 0;
}

int svo_ctx_last_pass_stats(svo_ctx* ctx, svo_pass_stats* out) {
  if (!ctx || !out) return SVO_E_INVALID;
  std::lock_guard<std::mutex> lk(ctx->impl->stage_mu);
  *out = ctx->impl->last_pass;
  return SVO_OK;
}

int svo_device_alloc(svo_ctx* ctx, uint64_t bytes, void** out) {
  if (!ctx || !out) return SVO_E_INVALID;
  return ctx->impl->dev_malloc(out, bytes);
}

int svo_device_free(svo_ctx* ctx, void* ptr, uint64_t bytes) {
  if (!ctx) return SVO_E_INVALID;
  ctx->impl->dev_free(ptr, bytes);
  return SVO_OK;
}

int svo_device_copy(svo_ctx* ctx, void* dst, const void* src, uint64_t bytes, int kind) {
  if (!ctx || kind < 0 || kind > 2) return SVO_E_INVALID;
  if (!bytes) return SVO_OK;
  Context& C = *ctx->impl;
  const hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
  std::lock_guard<std::mutex> lk(C.aux_mu);
  hipSetDevice(C.device);
  if (hipMemcpyAsync(dst, src, bytes, k, C.aux_stream) != hipSuccess) return SVO_E_HIP;
  return hipStreamSynchronize(C.aux_stream) == hipSuccess ? SVO_OK : SVO_E_HIP;
}

int svo_load_obj(const char* path, svo_mesh** out) {
  if (!path || !out) return SVO_E_INVALID;
  *out = nullptr;
  svo_mesh* m = new svo_mesh();
  const int st = load_obj_file(path, m->points, m->tris);
  if (st != SVO_OK) {
    delete m;
    return st;
  }
  *out = m;
  return SVO_OK;
}

int svo_mesh_view(const svo_mesh* mesh, svo_mesh_desc* out) {
  if (!mesh || !out) return SVO_E_INVALID;
  out->xyz = mesh->points.data();
  out->n_points = mesh->points.size() / 3;
  out->tri = mesh->tris.data();
  out->n_tris = mesh->tris.size() / 3;
  return SVO_OK;
}

void svo_mesh_free(svo_mesh* mesh) { delete mesh; }

int svo_validate(const svo_mesh_desc* mesh, const svo_params* params) {
  return validate_params(mesh, params, nullptr);
}

uint64_t svo_estimate_bytes(uint64_t n_points, uint64_t n_tris, const svo_params* params) {
  return params ? estimate_bytes(n_points, n_tris, params) : 0;
}

uint64_t svo_estimate_bytes_mesh(const svo_mesh_desc* mesh, const svo_params* params) {
  return mesh && params ? estimate_bytes_mesh(mesh, params) : 0;
}

static int build_common(svo_ctx* ctx, const svo_mesh_desc* mesh, const svo_params* params,
                        int on_device, svo_tree** out) {
  if (!ctx || !mesh || !params || !out) return SVO_E_INVALID;
  *out = nullptr;
  std::vector<MeshInput> in(1);
  in[0].desc = *mesh;
  in[0].params = *params;
  in[0].on_device = on_device != 0;
  std::vector<svo_tree*> trees;
  std::vector<int> st;
  const int s = run_batch(*ctx->impl, in, trees, st);
  if (s != SVO_OK) return s;
  if (st[0] != SVO_OK) return st[0];
  *out = trees[0];
  return SVO_OK;
}

int svo_build(svo_ctx* ctx, const svo_mesh_desc* mesh, const svo_params* params, svo_tree** out) {
  return build_common(ctx, mesh, params, 0, out);
}

int svo_build_device(svo_ctx* ctx, const svo_mesh_desc* mesh, const svo_params* params,
                     svo_tree** out) {
  return build_common(ctx, mesh, params, 1, out);
}

int svo_build_batch(svo_ctx* ctx, const svo_mesh_desc* meshes, const svo_params* params,
                    uint32_t n, int meshes_on_device, svo_tree** trees, int* statuses) {
  if (!ctx || (n && (!meshes || !params || !trees))) return SVO_E_INVALID;
  std::vector<MeshInput> in(n);
  for (uint32_t i = 0; i < n; ++i) {
    in[i].desc = meshes[i];
    in[i].params = params[i];
    in[i].on_device = meshes_on_device != 0;
  }
  std::vector<svo_tree*> tr;
  std::vector<int> st;
  const int s = run_batch(*ctx->impl, in, tr, st);
  // a failed pass of a split batch marks its own meshes; the others' trees
  // are returned
  int agg = s;
  for (uint32_t i = 0; i < n; ++i) {
    trees[i] = tr[i];
    if (statuses) statuses[i] = st[i];
    if (agg == SVO_OK && st[i] != SVO_OK) agg = st[i];
  }
  return agg;
}

// ---- batch plans (recorded once, replayed as one HIP graph)
struct svo_plan {
  std::unique_ptr<BatchPlan> impl;
};

int svo_plan_create(svo_ctx* ctx, const svo_mesh_desc* meshes, const svo_params* params,
                    uint32_t n, int meshes_on_device, svo_plan** out) {
  if (!ctx || !out || !n || !meshes || !params) return SVO_E_INVALID;
  *out = nullptr;
  std::vector<MeshInput> in(n);
  for (uint32_t i = 0; i < n; ++i) {
    in[i].desc = meshes[i];
    in[i].params = params[i];
    in[i].on_device = meshes_on_device != 0;
  }
  auto p = std::make_unique<svo_plan>();
  const int st = plan_create(*ctx->impl, in, p->impl);
  if (st != SVO_OK) return st;
  *out = p.release();
  return SVO_OK;
}

int svo_plan_run(svo_plan* plan, svo_tree** trees, int* statuses, int* replayed) {
  if (!plan || !trees) return SVO_E_INVALID;
  std::vector<svo_tree*> tr;
  std::vector<int> st;
  bool rp = false;
  const int s = plan_run(*plan->impl, tr, st, rp);
  if (replayed) *replayed = rp ? 1 : 0;
  int agg = s;
  for (size_t i = 0; i < plan->impl->in.size(); ++i) {
    trees[i] = i < tr.size() ? tr[i] : nullptr;
    const int si = i < st.size() ? st[i] : s;
    if (statuses) statuses[i] = si;
    if (agg == SVO_OK && si != SVO_OK) agg = si;
  }
  return agg;
}

int svo_plan_destroy(svo_plan* plan) {
  if (!plan) return SVO_E_INVALID;
  delete plan;
  return SVO_OK;
}

// ---- tuning search and persistence
namespace {
double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
}  // namespace

int svo_ctx_autotune(svo_ctx* ctx, const svo_mesh_desc* meshes, const svo_params* params,
                     uint32_t n, int meshes_on_device, uint32_t reps, svo_tune_result* out) {
  if (!ctx || !n || !meshes || !params) return SVO_E_INVALID;
  if (reps == 0) reps = 3;
  Context& C = *ctx->impl;
  Tuning base;
  {
    std::lock_guard<std::mutex> lk(C.mu);
    base = C.tuning;
  }
  std::vector<svo_tree*> trees(n, nullptr);
  std::vector<int> st(n, SVO_OK);
  // mean wall time of one build of the sample under tuning t
  auto trial = [&](const Tuning& t, double& ms) -> int {
    {
      std::lock_guard<std::mutex> lk(C.mu);
      C.tuning = t;
    }
    std::vector<double> times;
    for (uint32_t r = 0; r <= reps; ++r) {  // the first build warms up
      const double t0 = now_ms();
      const int rc = svo_build_batch(ctx, meshes, params, n, meshes_on_device, trees.data(),
                                     st.data());
      const double t1 = now_ms();
      for (auto& tr : trees) {
        svo_tree_free(tr);
        tr = nullptr;
      }
      if (rc != SVO_OK) return rc;
      if (r > 0) times.push_back(t1 - t0);
    }
    std::sort(times.begin(), times.end());
    ms = times[times.size() / 2];  // median: one slow build does not decide
    return SVO_OK;
  };
  double best_ms = 0, default_ms = 0;
  int rc = trial(base, default_ms);
  if (rc != SVO_OK) {
    std::lock_guard<std::mutex> lk(C.mu);
    C.tuning = base;
    return rc;
  }
  Tuning best = base;
  best_ms = default_ms;
  uint32_t trials = 1;
  const uint32_t smalls[] = {128, 256, 512, 1024, 2048};
  const uint32_t lanes[] = {1, 2, 3, 4};
  for (uint32_t sm : smalls)
    for (uint32_t ln : lanes) {
      if (ln > 1 && n < 2 * base.batch_min_meshes * (ln > 2 ? 2 : 1)) continue;
      Tuning t = base;
      t.small_max_cells = sm;
      t.batch_lanes = ln;
      if (t.small_max_cells == base.small_max_cells && t.batch_lanes == base.batch_lanes) continue;
      double ms = 0;
      rc = trial(t, ms);
      ++trials;
      if (rc != SVO_OK) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
  {
    std::lock_guard<std::mutex> lk(C.mu);
    C.tuning = rc == SVO_OK ? best : base;
  }
  // per size class: the class's meshes alone, timed on their voxelization
  std::vector<std::vector<uint32_t>> members(SIZE_CLASSES);
  for (uint32_t i = 0; i < n; ++i)
    members[size_class_of(params[i].res, meshes[i].n_tris)].push_back(i);
  const bool timing_was = C.timing;
  C.timing = true;
  auto vox_trial = [&](const std::vector<uint32_t>& sel, const Tuning& t, double& ms) -> int {
    {
      std::lock_guard<std::mutex> lk(C.mu);
      C.tuning = t;
    }
    std::vector<svo_mesh_desc> d;
    std::vector<svo_params> p;
    for (uint32_t i : sel) {
      d.push_back(meshes[i]);
      p.push_back(params[i]);
    }
    std::vector<svo_tree*> tr(sel.size(), nullptr);
    std::vector<int> s2(sel.size(), SVO_OK);
    std::vector<double> times;
    for (uint32_t r = 0; r <= reps; ++r) {
      const int rc2 = svo_build_batch(ctx, d.data(), p.data(), (uint32_t)sel.size(),
                                      meshes_on_device, tr.data(), s2.data());
      for (auto& x : tr) {
        svo_tree_free(x);
        x = nullptr;
      }
      if (rc2 != SVO_OK) return rc2;
      double v = 0;
      {
        std::lock_guard<std::mutex> lk(C.stage_mu);
        for (size_t k = 0; k < C.stage_names.size(); ++k)
          if (C.stage_names[k] == "classify" || C.stage_names[k] == "voxelize")
            v += C.stage_ms[k];
      }
      if (r > 0) times.push_back(v);
    }
    std::sort(times.begin(), times.end());
    ms = times[times.size() / 2];
    return SVO_OK;
  };
  const uint32_t cands[] = {64, 128, 256, 512, 1024, 2048, 4096};
  for (int c = 0; c < SIZE_CLASSES && rc == SVO_OK; ++c) {
    if (members[c].empty()) continue;
    Tuning t = best;
    t.class_small_max[c] = 0;
    double ref_ms = 0;
    rc = vox_trial(members[c], t, ref_ms);
    ++trials;
    double cls_best = ref_ms;
    uint32_t pick = 0;
    for (uint32_t v : cands) {
      if (rc != SVO_OK) break;
      if (v == best.small_max_cells) continue;
      t.class_small_max[c] = v;
      double ms = 0;
      rc = vox_trial(members[c], t, ms);
      ++trials;
      if (rc == SVO_OK && ms < cls_best) {
        cls_best = ms;
        pick = v;
      }
    }
    if (rc == SVO_OK && pick && cls_best < 0.97 * ref_ms) best.class_small_max[c] = pick;
  }
  C.timing = timing_was;
  {
    std::lock_guard<std::mutex> lk(C.mu);
    C.tuning = rc == SVO_OK ? best : base;
  }
  if (out) {
    out->small_max_cells = best.small_max_cells;
    out->batch_lanes = best.batch_lanes;
    out->best_ms = best_ms;
    out->default_ms = default_ms;
    out->n_trials = trials;
    for (int c = 0; c < SIZE_CLASSES; ++c) {
      out->class_small_max[c] = best.class_small_max[c];
      out->class_meshes[c] = (uint32_t)members[c].size();
    }
  }
  return rc;
}

int svo_ctx_save_tuning(svo_ctx* ctx, const char* path) {
  if (!ctx || !path) return SVO_E_INVALID;
  FILE* f = std::fopen(path, "w");
  if (!f) return SVO_E_IO;
  for (const OptionRef& o : kOptions) {
    int64_t v = 0;
    if (svo_ctx_get_option(ctx, o.name, &v) == SVO_OK)
      std::fprintf(f, "%s %lld\n", o.name, (long long)v);
  }
  return std::fclose(f) == 0 ? SVO_OK : SVO_E_IO;
}

int svo_ctx_load_tuning(svo_ctx* ctx, const char* path) {
  if (!ctx || !path) return SVO_E_INVALID;
  FILE* f = std::fopen(path, "r");
  if (!f) return SVO_E_IO;
  char name[64];
  long long v = 0;
  int rc = SVO_OK;
  while (rc == SVO_OK) {
    const int got = std::fscanf(f, "%63s %lld", name, &v);
    if (got == EOF) break;
    if (got != 2) {
      rc = SVO_E_PARSE;
      break;
    }
    rc = svo_ctx_set_option(ctx, name, (int64_t)v);
  }
  std::fclose(f);
  return rc;
}

int svo_tree_info_get(const svo_tree* tree, svo_tree_info* out) {
  if (!tree || !out) return SVO_E_INVALID;
  *out = tree->info;
  return SVO_OK;
}

uint64_t svo_tree_field_bytes(const svo_tree* tree, int field) {
  if (!tree) return 0;
  const uint64_t kb = tree->key_bits_out == 64 ? 8 : 4;
  const svo_tree_info& I = tree->info;
  switch (field) {
    case SVO_FIELD_VOXELS: return kb * I.n_voxels;
    case SVO_FIELD_CODE: return kb * I.n_nodes;
    case SVO_FIELD_LEVEL_NODES: return 8ull * (uint64_t)I.depth;
    case SVO_FIELD_NODES_AOS: return 168ull * I.n_nodes;
    default: {
      const FieldRange r = field_range(tree, field);
      return r.elem_bytes * r.count;
    }
  }
}

const void* svo_tree_device_ptr(const svo_tree* tree, int field) {
  if (!tree) return nullptr;
  return field_range(tree, field).dev;
}

int svo_tree_copy_field(const svo_tree* tree, int field, void* dst, uint64_t dst_bytes) {
  if (!tree || (!dst && svo_tree_field_bytes(tree, field))) return SVO_E_INVALID;
  if (field < 0 || field >= SVO_FIELD_COUNT) return SVO_E_INVALID;
  const uint64_t need = svo_tree_field_bytes(tree, field);
  if (dst_bytes < need) return SVO_E_RANGE;
  Context& C = *tree->ctx;
  if (field == SVO_FIELD_LEVEL_NODES) {
    std::memcpy(dst, tree->info.level_nodes, need);
    return SVO_OK;
  }
  if (field == SVO_FIELD_NODES_AOS) {
    // repack on the device in chunks, then copy
    const uint64_t N = tree->info.n_nodes;
    const uint64_t chunk = std::min<uint64_t>(N, 1ull << 20);
    if (!N) return SVO_OK;
    void* tmp = nullptr;
    int st = C.dev_malloc(&tmp, 168 * chunk);
    if (st != SVO_OK) return st;
    for (uint64_t a = 0; a < N && st == SVO_OK; a += chunk) {
      const uint64_t k = std::min(chunk, N - a);
      {
        std::lock_guard<std::mutex> lk(C.aux_mu);
        hipSetDevice(C.device);
        launch_pack_aos(tree->out->nodes, tree->node_base + a, k, tmp, C.aux_stream);
      }
      st = copy_d2h(C, (char*)dst + 168 * a, tmp, 168 * k);
    }
    C.dev_free(tmp, 168 * chunk);
    return st;
  }
  const FieldRange r = field_range(tree, field);
  if (!r.is_code || (uint64_t)tree->out->code_bytes * 8 == (uint64_t)tree->key_bits_out)
    return copy_d2h(C, dst, r.dev, r.elem_bytes * r.count);
  // stored width differs from the requested one: convert on the host
  std::vector<uint8_t> tmp(r.elem_bytes * r.count);
  int st = copy_d2h(C, tmp.data(), r.dev, tmp.size());
  if (st != SVO_OK) return st;
  for (uint64_t i = 0; i < r.count; ++i) {
    uint64_t v = 0;
    std::memcpy(&v, tmp.data() + r.elem_bytes * i, r.elem_bytes);
    if (tree->key_bits_out == 64) ((uint64_t*)dst)[i] = v;
    else ((uint32_t*)dst)[i] = (uint32_t)v;
  }
  return SVO_OK;
}

int svo_tree_copy_voxels(const svo_tree* tree, void* dst, uint64_t dst_bytes) {
  return svo_tree_copy_field(tree, SVO_FIELD_VOXELS, dst, dst_bytes);
}
int svo_tree_copy_nodes(const svo_tree* tree, svo_node_rec* dst, uint64_t dst_bytes) {
  return svo_tree_copy_field(tree, SVO_FIELD_NODES_AOS, dst, dst_bytes);
}
int svo_tree_copy_neighbors(const svo_tree* tree, uint32_t* dst, uint64_t dst_bytes) {
  return svo_tree_copy_field(tree, SVO_FIELD_NEIGHBORS, dst, dst_bytes);
}
int svo_tree_copy_vertices(const svo_tree* tree, svo_vertex* dst, uint64_t dst_bytes) {
  return svo_tree_copy_field(tree, SVO_FIELD_VERTICES, dst, dst_bytes);
}
int svo_tree_copy_edges(const svo_tree* tree, svo_edge* dst, uint64_t dst_bytes) {
  return svo_tree_copy_field(tree, SVO_FIELD_EDGES, dst, dst_bytes);
}
int svo_tree_copy_faces(const svo_tree* tree, svo_edge* dst, uint64_t dst_bytes) {
  return svo_tree_copy_field(tree, SVO_FIELD_FACES, dst, dst_bytes);
}

int svo_write_obj(const svo_tree* tree, const char* dir, const char* stem, char* path_out,
                  size_t path_cap) {
  if (!tree || !stem) return SVO_E_INVALID;
  // base_filename + "_" + treeDepth + "_tree.obj", SVO.cu:738-739
  std::string path = (dir && *dir) ? std::string(dir) + "/" : std::string();
  path += std::string(stem) + "_" + std::to_string(tree->info.depth) + "_tree.obj";
  if (path_out && path_cap) {
    std::strncpy(path_out, path.c_str(), path_cap - 1);
    path_out[path_cap - 1] = 0;
  }
  const uint64_t N = tree->info.n_nodes;
  std::vector<float> org(3 * N), w(N);
  int st = svo_tree_copy_field(tree, SVO_FIELD_ORIGIN, org.data(), 12 * N);
  if (st == SVO_OK) st = svo_tree_copy_field(tree, SVO_FIELD_WIDTH, w.data(), 4 * N);
  if (st != SVO_OK) return st;
  return write_tree_obj(path.c_str(), org.data(), w.data(), N);
}

int svo_write_voxels_obj(const svo_tree* tree, const char* dir, const char* stem,
                         char* path_out, size_t path_cap) {
  if (!tree || !stem) return SVO_E_INVALID;
  // base_filename + "_" + treeDepth + "_voxel.obj" with treeDepth == 1, SVO.cu:412, 757-758
  std::string path = (dir && *dir) ? std::string(dir) + "/" : std::string();
  path += std::string(stem) + "_1_voxel.obj";
  if (path_out && path_cap) {
    std::strncpy(path_out, path.c_str(), path_cap - 1);
    path_out[path_cap - 1] = 0;
  }
  const uint64_t U = tree->info.n_voxels;
  const int cb = tree->key_bits_out / 8;
  std::vector<uint8_t> raw(cb * U);
  int st = svo_tree_copy_field(tree, SVO_FIELD_VOXELS, raw.data(), raw.size());
  if (st != SVO_OK) return st;
  // modelBBox.min + width * Vector3f(x, y, z), width = unitNodeWidth, SVO.cu:769-770
  const float w = tree->info.unit[0];
  std::vector<float> org(3 * U), wid(U, w);
  for (uint64_t i = 0; i < U; ++i) {
    uint64_t code = 0;
    std::memcpy(&code, raw.data() + cb * i, cb);
    uint32_t c[3] = {0, 0, 0};
    for (int b = 0; b < 21; ++b)
      for (int a = 0; a < 3; ++a) c[a] |= (uint32_t)((code >> (3 * b + a)) & 1u) << b;
    for (int a = 0; a < 3; ++a) org[3 * i + a] = tree->info.bbox_min[a] + w * (float)c[a];
  }
  return write_tree_obj(path.c_str(), org.data(), wid.data(), U);
}

int svo_query_neighbors(const svo_tree* tree, const uint32_t* node, const uint8_t* slot,
                        uint64_t n, uint32_t* out) {
  if (!tree || (n && (!node || !slot || !out))) return SVO_E_INVALID;
  if (!n) return SVO_OK;
  Context& C = *tree->ctx;
  const size_t bytes = 4 * n + 4 * n + n + 512;
  void* tmp = nullptr;
  int st = C.dev_malloc(&tmp, bytes);
  if (st != SVO_OK) return st;
  uint32_t* d_node = (uint32_t*)tmp;
  uint32_t* d_out = d_node + n;
  uint8_t* d_slot = (uint8_t*)(d_out + n);
  {
    std::lock_guard<std::mutex> lk(C.aux_mu);
    hipSetDevice(C.device);
    hipStream_t s = C.aux_stream;
    if (hipMemcpyAsync(d_node, node, 4 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_slot, slot, n, hipMemcpyHostToDevice, s) != hipSuccess) {
      st = SVO_E_HIP;
    } else {
      launch_query_neighbors(tree->out->nodes.neighbors + 27 * tree->node_base,
                             tree->info.n_nodes, d_node, d_slot, n, d_out, s);
      if (hipMemcpyAsync(out, d_out, 4 * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        st = SVO_E_HIP;
    }
  }
  C.dev_free(tmp, bytes);
  return st;
}

void svo_tree_free(svo_tree* tree) { delete tree; }

}  // extern "C"
