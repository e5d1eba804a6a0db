// pipeline.cpp -- one pass of the SVO build over a batch of meshes.
//
// Host C++ drives the gfx950 kernels on one stream (a "lane"):
//
//   A0  bbox -> grid params -> triangle classification     (1 readback: box-cell
//                                                            total, big chunks)
//   A1  voxelize -> radix sort -> unique + level 1 -> levels 2..L
//                                                         (1 readback: per-mesh
//                                                          depth / level spans)
//   B   node records -> roots -> voxels -> neighbors -> vertex/edge/face
//                                                         (1 readback: totals)
//
// The readbacks size the next phase's buffers exactly; inside a phase every
// launch shape is fixed by host-known bounds and every data-dependent count
// is read by the kernels from device memory.  A single mesh is a batch of one;
// a batch sorts composite (mesh, Morton) keys, so all meshes share every
// launch.  Reference correspondence: SparseVoxelOctree::create (SVO.cu:341-520)
// and constructNodeAtrributes (SVO.cu:730-735).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include <thread>

#include "svo_host.h"

namespace svo {

#define HIP_TRY(expr)                  \
  do {                                 \
    hipError_t e_ = (expr);            \
    if (e_ != hipSuccess) return SVO_E_HIP; \
  } while (0)
#define SVO_TRY(expr)       \
  do {                      \
    int s_ = (expr);        \
    if (s_ != SVO_OK) return s_; \
  } while (0)

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ------------------------------------------------------------------ context
// Live contexts of the process: each caches freed blocks up to 80% of its
// device, so an allocation that fails in one context first makes the others
// on that device give their cached blocks back.
static std::mutex g_contexts_mu;
static std::vector<Context*> g_contexts;

static void trim_peer_caches(const Context* self) {
  std::lock_guard<std::mutex> lk(g_contexts_mu);
  for (Context* c : g_contexts)
    if (c != self && c->device == self->device) c->trim_cache(0);
}

// Lane streams.  A HIP stream is bound at creation to one of the device's
// hardware queues (GPU_MAX_HW_QUEUES of them, 4 by default); streams made
// beyond that share queues, and the work of streams on one queue runs one
// after the other.  The concurrent passes of a batch overlap only when their
// lanes' streams sit on different queues, which holds for the first streams
// a process makes but not for streams made after others were destroyed (a
// second context's four lanes were measured on two queues, pairwise
// serialized).  So lane streams are made once per process and device, the
// first request making a burst of LANE_BURST main streams and then their
// side streams, before any other stream of the library (the first context
// on the device makes the burst before its own copy stream), and contexts
// borrow them: a context gives its lanes' streams back when it releases its
// lanes, and the next context reuses them with their queues.
namespace {
constexpr int LANE_BURST = 4;
struct PooledStreams {
  hipStream_t stream, side;
  bool busy;
};
std::mutex g_stream_pool_mu;
std::map<int, std::vector<PooledStreams>> g_stream_pool;

// Appends n (stream, side) pairs to the pool of the current device: the n
// main streams first, then the side streams.  Caller holds g_stream_pool_mu.
int grow_pool(std::vector<PooledStreams>& pool, int n) {
  std::vector<hipStream_t> made;
  for (int i = 0; i < 2 * n; ++i) {
    hipStream_t q = nullptr;
    if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t m : made) hipStreamDestroy(m);
      return SVO_E_HIP;
    }
    made.push_back(q);
  }
  for (int i = 0; i < n; ++i) pool.push_back({made[i], made[n + i], false});
  return SVO_OK;
}

// Makes the first burst of `dev` (the current device) if it has none yet.
int prepare_streams(int dev) {
  std::lock_guard<std::mutex> lk(g_stream_pool_mu);
  std::vector<PooledStreams>& pool = g_stream_pool[dev];
  return pool.empty() ? grow_pool(pool, LANE_BURST) : SVO_OK;
}

// Lends a (stream, side) pair of `dev` (the current device) to a lane.
int take_streams(int dev, hipStream_t& stream, hipStream_t& side) {
  std::lock_guard<std::mutex> lk(g_stream_pool_mu);
  std::vector<PooledStreams>& pool = g_stream_pool[dev];
  size_t i = 0;
  while (i < pool.size() && pool[i].busy) ++i;
  if (i == pool.size()) {
    const int st = grow_pool(pool, pool.empty() ? LANE_BURST : 1);
    if (st != SVO_OK) return st;
  }
  pool[i].busy = true;
  stream = pool[i].stream;
  side = pool[i].side;
  return SVO_OK;
}

void give_streams(int dev, hipStream_t stream) {
  std::lock_guard<std::mutex> lk(g_stream_pool_mu);
  for (PooledStreams& p : g_stream_pool[dev])
    if (p.stream == stream) p.busy = false;
}
}  // namespace

int Context::init(int dev, uint64_t bud) {
  device = dev;
  budget = bud;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || dev < 0 || dev >= n) return SVO_E_HIP;
  HIP_TRY(hipSetDevice(dev));
  SVO_TRY(prepare_streams(dev));  // the lanes' streams take the first queues
  HIP_TRY(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
  // freed blocks stay cached up to 80% of the device (hipMalloc of a large
  // block maps its pages and costs seconds at 100+ GB)
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b) cache_limit = total_b / 10 * 8;
  std::lock_guard<std::mutex> lk(g_contexts_mu);
  g_contexts.push_back(this);
  return SVO_OK;
}

void Context::release_lanes() {
  std::lock_guard<std::mutex> lk(mu);
  hipSetDevice(device);
  for (auto& l : lanes) {
    if (l->stream) hipStreamSynchronize(l->stream);
    if (l->side) hipStreamSynchronize(l->side);
    l->ws.release_all();
    for (auto e : l->events) hipEventDestroy(e);
    if (l->fork) hipEventDestroy(l->fork);
    if (l->join) hipEventDestroy(l->join);
    if (l->pinned) hipHostFree(l->pinned);
    if (l->stream) give_streams(device, l->stream);
  }
  lanes.clear();
  free_lanes.clear();
  trim_cache(0);
}

Context::~Context() {
  {
    std::lock_guard<std::mutex> lk(g_contexts_mu);
    g_contexts.erase(std::remove(g_contexts.begin(), g_contexts.end(), this), g_contexts.end());
  }
  release_lanes();
  hipSetDevice(device);
  if (aux_stream) hipStreamDestroy(aux_stream);
  aux_stream = nullptr;
}

// A batch's sub-batch pass g asks for lane g: the sub-batches of a batch
// are the same from call to call, so each lane's workspaces fit its pass
// after the first call instead of growing (and allocating) whenever a pass
// lands on a lane that last ran a smaller one.  A busy preferred lane (other
// threads building on the context) yields any free one.
Lane* Context::acquire_lane(int prefer) {
  std::unique_lock<std::mutex> lk(mu);
  auto make_lane = [&]() -> bool {
    auto l = std::make_unique<Lane>();
    hipSetDevice(device);
    if (take_streams(device, l->stream, l->side) != SVO_OK) return false;
    if (hipEventCreateWithFlags(&l->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&l->join, hipEventDisableTiming) != hipSuccess) {
      give_streams(device, l->stream);
      if (l->fork) hipEventDestroy(l->fork);
      if (l->join) hipEventDestroy(l->join);
      return false;
    }
    l->ws.ctx = this;
    free_lanes.push_back(l.get());
    lanes.push_back(std::move(l));
    return true;
  };
  while (prefer >= 0 && (size_t)prefer >= lanes.size() && lanes.size() < MAX_LANES)
    if (!make_lane()) return nullptr;
  if (free_lanes.empty() && lanes.size() < MAX_LANES && !make_lane()) return nullptr;
  cv.wait(lk, [&] { return !free_lanes.empty(); });
  if (prefer >= 0 && (size_t)prefer < lanes.size()) {
    auto it = std::find(free_lanes.begin(), free_lanes.end(), lanes[prefer].get());
    if (it != free_lanes.end()) {
      Lane* l = *it;
      free_lanes.erase(it);
      return l;
    }
  }
  Lane* l = free_lanes.back();
  free_lanes.pop_back();
  return l;
}

void Context::release_lane(Lane* l) {
  {
    std::lock_guard<std::mutex> lk(mu);
    free_lanes.push_back(l);
  }
  cv.notify_one();
}

int Context::reserve(uint64_t bytes) {
  uint64_t cur = in_use.load();
  while (true) {
    if (budget && cur + bytes > budget) return SVO_E_OOM;
    if (in_use.compare_exchange_weak(cur, cur + bytes)) return SVO_OK;
  }
}
void Context::release(uint64_t bytes) { in_use.fetch_sub(bytes); }

// Size classes 2^k * {1, 1.25, 1.5, 1.75} above 1 MiB (at most 25% slack),
// 256-byte granules below.
size_t Context::size_class(size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (bytes <= (1u << 20)) return (bytes + 255) & ~(size_t)255;
  int k = 63 - __builtin_clzll((unsigned long long)bytes);
  const size_t base = (size_t)1 << k;
  const size_t step = base >> 2;
  return base + ((bytes - base + step - 1) / step) * step;
}

int Context::dev_malloc(void** p, size_t bytes) {
  *p = nullptr;
  const size_t cls = size_class(bytes);
  SVO_TRY(reserve(cls));
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    auto it = cache.find(cls);
    if (it != cache.end()) {
      *p = it->second;
      cache.erase(it);
      cached_bytes -= cls;
      return SVO_OK;
    }
  }
  hipSetDevice(device);
  if (hipMalloc(p, cls) != hipSuccess) {
    hipGetLastError();
    trim_cache(0);  // give cached blocks back and retry
    bool ok = hipMalloc(p, cls) == hipSuccess;
    if (!ok) {
      hipGetLastError();
      trim_peer_caches(this);  // then those of the other contexts on the device
      hipSetDevice(device);
      ok = hipMalloc(p, cls) == hipSuccess;
    }
    if (!ok) {
      hipGetLastError();
      release(cls);
      *p = nullptr;
      return SVO_E_OOM;
    }
  }
  n_mallocs.fetch_add(1);
  return SVO_OK;
}

void Context::dev_free(void* p, size_t bytes) {
  if (!p) return;
  const size_t cls = size_class(bytes);
  release(cls);
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    cache.emplace(cls, p);
    cached_bytes += cls;
  }
  uint64_t limit = cache_limit;
  if (budget) limit = std::min<uint64_t>(limit, budget / 2);
  if (cached_bytes > limit) trim_cache(limit);
}

void Context::trim_cache(uint64_t keep_bytes) {
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    // release the largest blocks first
    while (cached_bytes > keep_bytes && !cache.empty()) {
      auto it = std::prev(cache.end());
      cached_bytes -= it->first;
      drop.push_back(it->second);
      cache.erase(it);
    }
  }
  if (drop.empty()) return;
  hipSetDevice(device);
  for (void* q : drop) hipFree(q);
}

DeviceBlock::~DeviceBlock() {
  if (ctx && ptr) ctx->dev_free(ptr, bytes);
}

int Workspace::get(const char* name, size_t bytes, void** out) {
  bytes = align256(bytes ? bytes : 256);
  Buf& b = bufs[name];
  // grow when too small; give back a buffer far larger than this pass needs
  // (after one huge pass the lane would otherwise hold its memory for good)
  const bool oversize = b.bytes > 4 * bytes && b.bytes - bytes > (1ull << 30);
  if (frozen) {
    if (!b.ptr || b.bytes < bytes) return SVO_E_HIP;  // would allocate
    *out = b.ptr;
    return SVO_OK;
  }
  if (b.bytes < bytes || oversize) {
    if (b.ptr) {
      ctx->dev_free(b.ptr, b.bytes);
      ctx->ws_bytes.fetch_sub(Context::size_class(b.bytes));
    }
    b.ptr = nullptr;
    b.bytes = 0;
    b.clean = 0;
    // grow with headroom so that slowly growing inputs do not reallocate
    const size_t want = bytes + bytes / 8;
    void* p = nullptr;
    int st = ctx->dev_malloc(&p, want);
    if (st != SVO_OK) {
      st = ctx->dev_malloc(&p, bytes);
      if (st != SVO_OK) return st;
      b.bytes = bytes;
    } else {
      b.bytes = want;
    }
    b.ptr = p;
    ctx->ws_bytes.fetch_add(Context::size_class(b.bytes));  // as in_use counts it
  }
  *out = b.ptr;
  return SVO_OK;
}

void Workspace::release_all() {
  for (auto& kv : bufs)
    if (kv.second.ptr) {
      ctx->dev_free(kv.second.ptr, kv.second.bytes);
      ctx->ws_bytes.fetch_sub(Context::size_class(kv.second.bytes));
    }
  bufs.clear();
}

size_t Workspace::total() const {
  size_t s = 0;
  for (auto& kv : bufs) s += kv.second.bytes;
  return s;
}

int Lane::ensure_pinned(size_t bytes) {
  if (pinned_bytes >= bytes) return SVO_OK;
  if (pinned) hipHostFree(pinned);
  pinned = nullptr;
  pinned_bytes = 0;
  const size_t want = std::max<size_t>(bytes, 1 << 16);
  if (hipHostMalloc(&pinned, want, hipHostMallocDefault) != hipSuccess) return SVO_E_HIP;
  pinned_bytes = want;
  return SVO_OK;
}

// ------------------------------------------------------------------ validation
static int bits_for(int32_t r) {
  int b = 0;
  while ((1ll << b) < (long long)r) ++b;
  return b;
}

int validate_params(const svo_mesh_desc* mesh, const svo_params* p, int* key_bits_out) {
  if (!mesh || !p) return SVO_E_INVALID;
  if (mesh->n_points == 0 || mesh->n_tris == 0) return SVO_E_EMPTY;
  if (!mesh->xyz || !mesh->tri) return SVO_E_INVALID;
  if (p->semantics > SVO_SEM_INTENDED) return SVO_E_INVALID;
  if (mesh->n_points >= (1ull << 31) || mesh->n_tris >= (1ull << 32)) return SVO_E_INVALID;
  int kb = (int)p->key_bits;
  int32_t rmax = 0;
  for (int a = 0; a < 3; ++a) {
    if (p->res[a] < 1) return SVO_E_INVALID;
    rmax = std::max(rmax, p->res[a]);
  }
  if (kb == 0) kb = rmax <= 1024 ? 32 : 64;
  if (kb != 32 && kb != 64) return SVO_E_INVALID;
  // uint32 codes: bit 31 is the reference's flag and the 16-bit-coordinate
  // encoder drops bit 32 (MortonLUT.h:482-501, SVO.cu:180-182): res <= 1024.
  if (kb == 32 && rmax > 1024) return SVO_E_RES_LIMIT;
  // uint64 codes: coordinates stay within the reference's uint16 loop range.
  if (kb == 64 && rmax > 65536) return SVO_E_RES_LIMIT;
  if (key_bits_out) *key_bits_out = kb;
  return SVO_OK;
}

// Host-side device-memory estimate for a build: surface voxels ~ 4.7 res^2
// for a closed surface spanning the box, nodes ~ 2.7 per voxel, output
// ~630 B per node (records of the reference layout), plus key workspace.
uint64_t estimate_bytes(uint64_t n_points, uint64_t n_tris, const svo_params* p) {
  const double r = std::max({p->res[0], p->res[1], p->res[2]});
  const double voxels = std::min(4.7 * r * r * 2.0, r * r * r);
  const double hits = std::max(voxels * 3.0, (double)n_tris * 4.0);
  const double nodes = 2.7 * voxels;
  const double per_node = (p->flags & SVO_BUILD_NO_VERTEX_EDGE) ? 170.0
                          : (p->semantics == SVO_SEM_INTENDED ? 800.0 : 640.0);
  const double b = nodes * per_node + hits * 8.0 * 2.5 + (double)n_points * 12.0 +
                   (double)n_tris * 12.0 + (double)order_bytes(n_tris) + (1 << 20);
  return (uint64_t)b;
}

// Upper bounds on the device bytes one build takes on a fresh lane: its
// output block plus every workspace buffer of the pass, with the growth
// headroom and size-class rounding of the allocator, for either
// de-duplication path.  `U` are bounds on the level code counts U_1..U_L,
// `cells` on the candidate cells (box cells summed over the triangles) and
// `rows` on their box rows.
namespace {
double pass_bytes_bound(const svo_params* p, int kb, uint64_t P, uint64_t T, int levels,
                        const std::vector<double>& U, double cells, double rows) {
  auto cells_at = [&](int l) {
    double c = 1.0;
    for (int a = 0; a < 3; ++a)
      c *= (double)(((int64_t)p->res[a] + (1ll << (l - 1)) - 1) >> (l - 1));
    return c;
  };
  const double K = kb == 64 ? 8.0 : 4.0;
  double N = 1.0;
  for (int l = 2; l <= levels; ++l) N += 8.0 * U[l];
  // output block (run_typed's layout)
  const bool attrs = !(p->flags & SVO_BUILD_NO_VERTEX_EDGE);
  const double rec =
      attrs ? (16.0 * 8 + 28.0 * 12 + (p->semantics == SVO_SEM_INTENDED ? 28.0 * 6 : 0)) : 0.0;
  const double out = (K + 161.0 + rec) * N + K * U[1] + 12 * 256.0;
  // workspace
  const uint64_t bm_words = bitmap_region_words(
      (uint64_t)morton3<uint64_t>((uint32_t)p->res[0] - 1, (uint32_t)p->res[1] - 1,
                                  (uint32_t)p->res[2] - 1) + 1);
  const double key_cap = std::max(cells, 1.0);
  const uint64_t kc = (uint64_t)std::min(key_cap, 4.0e18);
  double lv = 0.0, cap = key_cap;
  for (int l = 1; l <= levels + 1; ++l) {
    if (l >= 2) cap = std::min(cap, cells_at(l));
    lv += K * cap + (l >= 2 ? 8.0 * cap : 0.0) + 3 * 256.0;
  }
  const double n_big = (double)T + rows / 512.0;
  const double ws_common = 12.0 * P + 12.0 * T + (double)order_bytes(T) + 2 * 256.0 + lv +
                           (double)level_tmp_bytes(kc) +
                           sizeof(BigItem) * std::max(n_big, 1.0) + (double)attr_tmp_bytes(4096) +
                           (1 << 20);  // control buffers of one mesh
  const double ws_bitmap = 4.0 * bm_words + (double)bitmap_level_tmp_bytes(bm_words);
  const double ws_sort = 2.0 * K * key_cap + (double)radix_sort_tmp_bytes(kc);
  const double ws = ws_common + std::max(ws_bitmap, ws_sort);
  // allocator: 1/8 growth headroom on workspace buffers, size classes <= 1.25x
  return 1.25 * out + 1.25 * 1.125 * ws + (2 << 20);
}

int levels_of(const svo_params* p) {
  int levels = 1;  // levels until every axis has one cell
  while (true) {
    bool one = true;
    for (int a = 0; a < 3; ++a)
      if (((int64_t)p->res[a] + (1ll << (levels - 1)) - 1) >> (levels - 1) > 1) one = false;
    if (one) return levels;
    ++levels;
  }
}

double cells_of_level(const svo_params* p, int l) {
  double c = 1.0;
  for (int a = 0; a < 3; ++a) c *= (double)(((int64_t)p->res[a] + (1ll << (l - 1)) - 1) >> (l - 1));
  return c;
}

// Morton code of coordinates below 2^16 by byte lookup (same value as morton3)
struct MortonLut {
  uint32_t spread[256];
  MortonLut() {
    for (uint32_t b = 0; b < 256; ++b) {
      uint32_t v = 0;
      for (int i = 0; i < 8; ++i) v |= ((b >> i) & 1u) << (3 * i);
      spread[b] = v;
    }
  }
  uint64_t sp(uint32_t c) const {
    return (uint64_t)spread[c & 255] | ((uint64_t)spread[(c >> 8) & 255] << 24);
  }
  uint64_t operator()(uint32_t x, uint32_t y, uint32_t z) const {
    return sp(x) | (sp(y) << 1) | (sp(z) << 2);
  }
};
const MortonLut g_morton_lut;
}  // namespace

// The bound from the grid alone: every level full (U_l = cells of level l),
// and candidate cells no more than the automatic path choice allows on the
// sort path (it sorts only when the bitmap is over 64 bytes per candidate
// cell; the bitmap path keeps at most one key per grid cell).  Not a bound
// when a context forces the sort path.
uint64_t grid_bound_bytes(const svo_params* p, int kb, uint64_t n_points, uint64_t n_tris) {
  const int levels = levels_of(p);
  std::vector<double> U(levels + 2, 0.0);
  for (int l = 1; l <= levels; ++l) U[l] = cells_of_level(p, l);
  const double bm_bytes = 4.0 * (double)(((uint64_t)morton3<uint64_t>(
                                              (uint32_t)p->res[0] - 1, (uint32_t)p->res[1] - 1,
                                              (uint32_t)p->res[2] - 1) + 1 + 31) / 32);
  const double rows = (double)n_tris * (double)p->res[1] * (double)p->res[2];
  const double b = pass_bytes_bound(p, kb, n_points, n_tris, levels, U,
                                    std::min(U[1], bm_bytes / 64.0), rows);
  return b >= 1.8e19 ? UINT64_MAX : (uint64_t)b;
}

// The bound from the mesh: exact where the pass's sizes follow from the
// triangles alone (the grid, each triangle's box of candidate cells, hence
// the key capacity and the bitmap size); for the data-dependent counts
//   voxels U_1 <= min(candidate cells, grid cells, sum over triangles of the
//     cells their plane can cross: a triangle whose normal in cell units m
//     has dominant axis d crosses at most floor((|m_a| + |m_b|) / |m_d|) + 2
//     cells of a column along d, one more here for rounding),
//   level-l codes U_l <= min(U_(l-1), grid cells of level l, that per-triangle
//     sum at level l for the fine levels, and from a level k on the union of
//     the level-k cells that the triangles' planes can touch inside their
//     boxes),
//   nodes N <= 1 + 8 sum_(l>=2) U_l.
// A hit voxel lies in its triangle's box and touches its plane, so its level-k
// ancestor touches the plane inside the box's level-k image.  The union is
// counted on a level-k occupancy bitmap in Morton order, whose nonzero bytes
// are the level-(k+1) union, and so on up; k is the finest level at which the
// triangles' boxes hold few enough cells that marking stays cheap.
// Falls back to the count-only estimate for inputs the build rejects.
uint64_t estimate_bytes_mesh(const svo_mesh_desc* m, const svo_params* p) {
  int kb = 32;
  if (validate_params(m, p, &kb) != SVO_OK)
    return m && p ? estimate_bytes(m->n_points, m->n_tris, p) : 0;
  const uint64_t P = m->n_points, T = m->n_tris;
  // grid: createMeshBBCube (BaseModel.h:24-49) and the voxel size
  // (SVO.cu:348-350), in the same float operations as k_params
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = 0; i < P; ++i)
    for (int a = 0; a < 3; ++a) {
      const float v = m->xyz[3 * i + a];
      if (!std::isfinite(v)) return estimate_bytes(P, T, p);
      mn[a] = fmin_r(v, mn[a]);
      mx[a] = fmax_r(v, mx[a]);
    }
  float len[3], lo[3], hi[3], unit[3];
  for (int a = 0; a < 3; ++a) {
    len[a] = mx[a] - mn[a];
    lo[a] = mn[a];
    hi[a] = mx[a];
  }
  const float L = fmax_r(len[0], fmax_r(len[1], len[2]));
  if (!(L > 0.0f)) return estimate_bytes(P, T, p);
  for (int a = 0; a < 3; ++a)
    if (L != len[a]) {
      const float delta = L - len[a];
      lo[a] = mn[a] - (delta / 2.0f);
      hi[a] = mx[a] + (delta / 2.0f);
    }
  for (int a = 0; a < 3; ++a) {
    const float eps = (hi[a] - lo[a]) / 10001.0f;
    lo[a] -= eps;
    hi[a] += eps;
    unit[a] = (hi[a] - lo[a]) / (float)p->res[a];
  }
  // Each point in cells: (p - gridOrigin) / unit, rounded as the voxelizer
  // rounds it.  Rounded division by a positive unit is monotone, so the
  // voxelizer's (int)(min_k v_k / unit) is (int)(min_k q_k): the triangle
  // boxes below are the voxelizer's exactly (SVO.cu:73-85).
  std::vector<float> q(3 * P);
  for (uint64_t i = 0; i < P; ++i)
    for (int a = 0; a < 3; ++a) q[3 * i + a] = (m->xyz[3 * i + a] - lo[a]) / unit[a];
  const int levels = levels_of(p);
  const int rmax[3] = {p->res[0] - 1, p->res[1] - 1, p->res[2] - 1};
  // pass 1: boxes, candidate cells and rows, box cells per level
  std::vector<int32_t> box(6 * T);
  std::vector<double> boxsum(levels + 2, 0.0);
  double cells = 0.0, rows = 0.0;
  for (uint64_t t = 0; t < T; ++t) {
    const int32_t* ti = m->tri + 3 * t;
    if (ti[0] < 0 || ti[1] < 0 || ti[2] < 0 || (uint64_t)ti[0] >= P || (uint64_t)ti[1] >= P ||
        (uint64_t)ti[2] >= P)
      return estimate_bytes(P, T, p);  // the build rejects the mesh
    const float* v0 = &q[3 * (uint64_t)ti[0]];
    const float* v1 = &q[3 * (uint64_t)ti[1]];
    const float* v2 = &q[3 * (uint64_t)ti[2]];
    int32_t* g = &box[6 * t];
    for (int a = 0; a < 3; ++a) {
      g[a] = clamp_i((int)fmin_r(v0[a], fmin_r(v1[a], v2[a])), 0, rmax[a]);
      g[3 + a] = clamp_i((int)fmax_r(v0[a], fmax_r(v1[a], v2[a])), 0, rmax[a]);
    }
    const double e1 = (double)(g[3] - g[0] + 1) * (g[4] - g[1] + 1) * (g[5] - g[2] + 1);
    cells += e1;
    rows += (double)(g[4] - g[1] + 1) * (g[5] - g[2] + 1);
    boxsum[1] += e1;
    for (int l = 2; l <= levels; ++l) {
      const int sh = l - 1;
      boxsum[l] += (double)((g[3] >> sh) - (g[0] >> sh) + 1) * ((g[4] >> sh) - (g[1] >> sh) + 1) *
                   ((g[5] >> sh) - (g[2] >> sh) + 1);
    }
  }
  // union level: the finest level >= 2 whose boxes hold at most 32 cells per
  // triangle on average (or 4M cells in all), and whose bitmap fits the byte
  // lookup and 256 MB
  int k = 0;
  uint64_t bits_k = 0;
  for (int l = 2; l <= levels && !k; ++l) {
    if (boxsum[l] > std::max(32.0 * (double)T, 4194304.0)) continue;
    uint32_t r[3];
    for (int a = 0; a < 3; ++a) r[a] = (uint32_t)(((int64_t)p->res[a] + (1ll << (l - 1)) - 1) >> (l - 1));
    if (std::max({r[0], r[1], r[2]}) > 65536u) continue;
    const uint64_t bits = morton3<uint64_t>(r[0] - 1, r[1] - 1, r[2] - 1) + 1;
    if (bits > (1ull << 31)) continue;
    k = l;
    bits_k = bits;
  }
  const int fine_end = k ? k - 1 : levels;  // per-triangle sums on levels 1..fine_end
  std::vector<uint8_t> occ;
  if (k) occ.assign((size_t)((bits_k + 7) / 8 + 8), 0);
  const int ks = k ? k - 1 : 0;  // level-k cell = voxel >> ks
  const double inv_s = 1.0 / (double)(1ll << ks);
  std::vector<double> lvl_bound(levels + 2, INFINITY);
  for (int l = 1; l <= fine_end; ++l) lvl_bound[l] = 0.0;
  // pass 2: plane limits on the fine levels, marks on level k
  for (uint64_t t = 0; t < T; ++t) {
    const int32_t* g = &box[6 * t];
    const int ex[3] = {g[3] - g[0], g[4] - g[1], g[5] - g[2]};
    bool slab = false;
    int d = 0, a1 = 1, a2 = 2;
    double per_col = INFINITY, ga = 0.0, gb = 0.0, z0 = 0.0, x0 = 0.0, y0 = 0.0;
    if (ex[0] > 2 || ex[1] > 2 || ex[2] > 2) {
      // plane normal in cell units (boxes of at most 3 cells per axis are
      // counted whole: their plane limits little)
      const int32_t* ti = m->tri + 3 * t;
      const float* v0 = &q[3 * (uint64_t)ti[0]];
      const float* v1 = &q[3 * (uint64_t)ti[1]];
      const float* v2 = &q[3 * (uint64_t)ti[2]];
      const double e0[3] = {(double)v1[0] - v0[0], (double)v1[1] - v0[1], (double)v1[2] - v0[2]};
      const double e1[3] = {(double)v2[0] - v1[0], (double)v2[1] - v1[1], (double)v2[2] - v1[2]};
      const double mm[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                            e0[0] * e1[1] - e0[1] * e1[0]};
      const double am[3] = {std::fabs(mm[0]), std::fabs(mm[1]), std::fabs(mm[2])};
      d = am[0] >= am[1] ? (am[0] >= am[2] ? 0 : 2) : (am[1] >= am[2] ? 1 : 2);
      a1 = d == 2 ? 0 : d + 1;
      a2 = d == 0 ? 2 : (d == 1 ? 0 : 1);
      // the voxelizer's float normal can point elsewhere on a sliver (edges
      // nearly parallel, |e0 x e1| < 1e-3 |e0| |e1|): no plane limit then
      const double l0 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
      const double l1 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
      const double lc = am[0] * am[0] + am[1] * am[1] + am[2] * am[2];
      slab = am[d] > 0.0 && lc > 1e-6 * l0 * l1;
      if (slab) {
        const double inv_d = 1.0 / mm[d];
        per_col = std::floor((am[a1] + am[a2]) * std::fabs(inv_d)) + 3.0;
        ga = -mm[a1] * inv_d;
        gb = -mm[a2] * inv_d;
        z0 = v0[d] * inv_s;
        x0 = v0[a1] * inv_s;
        y0 = v0[a2] * inv_s;
      }
    }
    for (int l = 1; l <= fine_end; ++l) {
      const int sh = l - 1;
      const double e[3] = {(double)((g[3] >> sh) - (g[0] >> sh) + 1),
                           (double)((g[4] >> sh) - (g[1] >> sh) + 1),
                           (double)((g[5] >> sh) - (g[2] >> sh) + 1)};
      lvl_bound[l] += std::min(e[0] * e[1] * e[2], e[a1] * e[a2] * std::min(e[d], per_col));
    }
    if (!k) continue;
    int b0[3], b1[3];
    for (int a = 0; a < 3; ++a) {
      b0[a] = g[a] >> ks;
      b1[a] = g[3 + a] >> ks;
    }
    uint32_t c[3];
    for (int j = b0[a2]; j <= b1[a2]; ++j) {
      c[a2] = (uint32_t)j;
      for (int i = b0[a1]; i <= b1[a1]; ++i) {
        c[a1] = (uint32_t)i;
        int k0 = b0[d], k1 = b1[d];
        if (slab) {
          // the level-k cells of this column that the plane can touch: its
          // range over the column's corners, widened by a quarter cell for
          // rounding (the voxelizer's float plane differs from this one by
          // far less on a triangle that is not a sliver)
          const double zc = z0 + ga * (i - x0) + gb * (j - y0);
          const double zl = zc + std::min(0.0, ga) + std::min(0.0, gb);
          const double zh = zc + std::max(0.0, ga) + std::max(0.0, gb);
          k0 = (int)std::max((double)b0[d], std::floor(zl - 0.25));
          k1 = (int)std::min((double)b1[d], std::floor(zh + 0.25));
        }
        for (int kk = k0; kk <= k1; ++kk) {
          c[d] = (uint32_t)kk;
          const uint64_t code = g_morton_lut(c[0], c[1], c[2]);
          occ[code >> 3] |= (uint8_t)(1u << (code & 7));
        }
      }
    }
  }
  // union counts of levels k, k+1, ...: set bits, then each 8-byte word's
  // nonzero bytes as one byte of the next level
  for (int l = k; k && l <= levels; ++l) {
    const size_t nw = (occ.size() + 7) / 8;
    occ.resize(8 * nw, 0);
    uint64_t set = 0;
    std::vector<uint8_t> up(nw, 0);
    for (size_t i = 0; i < nw; ++i) {
      uint64_t w;
      std::memcpy(&w, occ.data() + 8 * i, 8);
      set += (uint64_t)__builtin_popcountll(w);
      uint64_t tt = w | (w >> 4);
      tt |= tt >> 2;
      tt |= tt >> 1;
      tt &= 0x0101010101010101ull;
      up[i] = (uint8_t)((tt * 0x0102040810204080ull) >> 56);
    }
    lvl_bound[l] = (double)set;
    occ.swap(up);
  }
  std::vector<double> U(levels + 2, 0.0);
  U[1] = std::min({std::max(cells, 1.0), cells_of_level(p, 1), lvl_bound[1]});
  for (int l = 2; l <= levels; ++l)
    U[l] = std::min({U[l - 1], cells_of_level(p, l), lvl_bound[l]});
  const double b = pass_bytes_bound(p, kb, P, T, levels, U, cells, rows);
  return b >= 1.8e19 ? UINT64_MAX : (uint64_t)b;
}

// ------------------------------------------------------------------ timing
// Figures of one svo_build* call, which may run as several concurrent
// passes: each stage's window runs from its earliest start to its latest end
// over the passes (times from a reference event recorded when the call
// began), and the pass sizes are summed, so that bytes / window is the
// stage's aggregate rate.
struct CallTiming {
  hipEvent_t ref = nullptr;
  std::mutex mu;
  std::vector<std::string> names;
  std::vector<double> t0, t1;  // per stage window, ms after ref
  svo_pass_stats ps{};
  bool any = false;
  void add(const std::vector<std::string>& nm, const std::vector<double>& a,
           const std::vector<double>& b, const svo_pass_stats& p) {
    std::lock_guard<std::mutex> lk(mu);
    for (size_t i = 0; i < nm.size(); ++i) {
      size_t k = 0;
      while (k < names.size() && names[k] != nm[i]) ++k;
      if (k == names.size()) {
        names.push_back(nm[i]);
        t0.push_back(a[i]);
        t1.push_back(b[i]);
      } else {
        t0[k] = std::min(t0[k], a[i]);
        t1[k] = std::max(t1[k], b[i]);
      }
    }
    if (!any) {
      ps = p;
    } else {
      ps.n_meshes += p.n_meshes;
      ps.n_tris += p.n_tris;
      ps.box_cells += p.box_cells;
      ps.n_hits += p.n_hits;
      ps.n_voxels += p.n_voxels;
      ps.level_elems += p.level_elems;
      ps.n_nodes += p.n_nodes;
      ps.n_vertices += p.n_vertices;
      ps.n_edges += p.n_edges;
      ps.n_faces += p.n_faces;
      ps.out_bytes += p.out_bytes;
      ps.bitmap_bytes += p.bitmap_bytes;
      ps.n_groups1 += p.n_groups1;
      ps.bitmap_fill_bytes += p.bitmap_fill_bytes;
      ps.bitmap_reset_bytes += p.bitmap_reset_bytes;
      ps.key_bytes = std::max(ps.key_bytes, p.key_bytes);
      ps.sort_bits = std::max(ps.sort_bits, p.sort_bits);
      ps.sort_passes = std::max(ps.sort_passes, p.sort_passes);
      ps.n_levels = std::max(ps.n_levels, p.n_levels);
    }
    any = true;
  }
  void publish(Context& C) {
    if (!any) return;
    std::lock_guard<std::mutex> lk(C.stage_mu);
    C.last_pass = ps;
    if (C.timing && !names.empty()) {
      C.stage_names = names;
      C.stage_ms.resize(names.size());
      for (size_t k = 0; k < names.size(); ++k) C.stage_ms[k] = t1[k] - t0[k];
    }
  }
};

struct StageTimer {
  Context& C;
  Lane& L;
  std::vector<std::string> names;
  size_t used = 0;
  // stages timed on the side stream: name, start event, end event (indices
  // into L.events past the main-stream marks)
  struct Range {
    std::string name;
    size_t a, b;
  };
  std::vector<Range> ranges;
  std::vector<hipEvent_t> side_ev;
  bool off = false;  // no events (a pass under graph capture)
  StageTimer(Context& c, Lane& l, bool disabled = false) : C(c), L(l), off(disabled) {}
  // Timing-only events: no system-scope fence when they complete (a fence
  // writes back and invalidates the caches at every stage boundary, which
  // lengthens the stage after it); the host reads them only after the pass's
  // streams are synchronized.
  static hipError_t create_event(hipEvent_t* e) {
    return hipEventCreateWithFlags(e, hipEventDisableSystemFence);
  }
  void mark(const char* name) {
    if (!C.timing || off) return;
    if (L.events.size() <= used) {
      hipEvent_t e;
      if (create_event(&e) != hipSuccess) return;
      L.events.push_back(e);
    }
    hipEventRecord(L.events[used], L.stream);
    names.push_back(name);
    ++used;
  }
  // Records a timing event on `st` (the side stream); returns its index.
  size_t side_mark(hipStream_t st) {
    if (!C.timing || off) return 0;
    const size_t i = side_ev.size();
    // side events live in L.events after a fixed block of 64 main marks
    const size_t slot = 64 + i;
    while (L.events.size() <= slot) {
      hipEvent_t e;
      if (create_event(&e) != hipSuccess) return 0;
      L.events.push_back(e);
    }
    hipEventRecord(L.events[slot], st);
    side_ev.push_back(L.events[slot]);
    return i;
  }
  void side_range(const char* name, size_t a, size_t b) {
    if (C.timing) ranges.push_back({name, a, b});
  }
  // Hands the pass's stage windows and sizes to the call's figures.
  void publish(CallTiming& ct, const svo_pass_stats& ps) {
    std::vector<double> a, b;
    std::vector<std::string> nm;
    if (C.timing && ct.ref && used >= 2)
      for (size_t i = 1; i < used; ++i) {
        float x = 0, y = 0;
        hipEventElapsedTime(&x, ct.ref, L.events[i - 1]);
        hipEventElapsedTime(&y, ct.ref, L.events[i]);
        a.push_back(x);
        b.push_back(y);
        nm.push_back(names[i]);
      }
    if (C.timing && ct.ref)
      for (const Range& r : ranges) {
        float x = 0, y = 0;
        hipEventElapsedTime(&x, ct.ref, side_ev[r.a]);
        hipEventElapsedTime(&y, ct.ref, side_ev[r.b]);
        a.push_back(x);
        b.push_back(y);
        nm.push_back(r.name);
      }
    ct.add(nm, a, b, ps);
  }
};

// ------------------------------------------------------------------ the pass
namespace {

// True when the reference's consecutive unique (uniqueVert / uniqueEdge,
// SVO.cu:679-720) provably removes nothing: every node width w is positive,
// finite and at least 2^-22 times the largest coordinate magnitude a node
// corner can have, so o + w != o on every axis.  Then the eight corners
// w*(bits) of a node are pairwise distinct, corner 7 of a node differs from
// corner 0 of the next, all eight absolute corners of a node are distinct,
// adjacent edges of a node differ, and edge 11 of a node (along z) cannot
// equal edge 0 of the next (along y).
bool no_adjacent_duplicates(const std::vector<MeshNodeLayout>& lay, const MeshParams* par,
                            const std::vector<int>& mstat) {
  for (size_t i = 0; i < lay.size(); ++i) {
    if (mstat[i] != SVO_OK || lay[i].n_nodes == 0) continue;
    for (uint32_t l = 0; l < lay[i].depth; ++l) {
      const double w = lay[i].width[l];
      if (!(w > 0.0) || !std::isfinite(w)) return false;
      // node origins are gridOrigin + w * coord with coord < ceil(res / 2^(l-1))
      // on every axis (the width is unitVoxelSize.x on all axes, SVO.cu:351)
      double M = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double cells = std::ceil((double)par[i].res[a] / std::ldexp(1.0, (int)l));
        M = std::max(M, std::fabs((double)par[i].bmin[a]) + w * (cells + 2.0));
      }
      if (w < M * 0x1p-22) return false;
    }
  }
  return true;
}


template <typename K>
int run_typed(Context& C, Lane& L, const std::vector<MeshInput>& in, const Plan& plan,
              std::vector<svo_tree*>& trees, std::vector<int>& statuses) {
  hipStream_t s = L.stream;
  Workspace& W = L.ws;
  const uint32_t n = (uint32_t)plan.idx.size();
  const uint32_t NL = (uint32_t)plan.n_levels;
  // A captured pass reads nothing back: the host values come from the
  // recorded run, the device compares its own with them (rec->d_flag), and
  // the buffers uploaded by the recorded run are still in place.
  PassRecord* rec = plan.rec;
  const bool capture = rec && rec->mode == PassRecord::CAPTURE;
  StageTimer timer(C, L, capture);
  KeyLayout<K> kl;
  kl.cb = plan.cb;
  kl.cmask = (plan.cb >= (int)(8 * sizeof(K))) ? (K)~(K)0 : (K)(((K)1 << plan.cb) - 1);

  // ---- inputs on the device
  std::vector<MeshDev> md(n);
  std::vector<uint64_t> tri_base(n + 1, 0);
  uint64_t host_bytes = 0, max_points = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const MeshInput& mi = in[plan.idx[i]];
    if (!mi.on_device) host_bytes += align256(12 * mi.desc.n_points) + align256(12 * mi.desc.n_tris);
  }
  char* staging = nullptr;
  if (host_bytes && capture) return SVO_E_INVALID;  // plans hold device inputs
  if (host_bytes) SVO_TRY(W.get("input", host_bytes, (void**)&staging));
  uint64_t soff = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const MeshInput& mi = in[plan.idx[i]];
    MeshDev& d = md[i];
    d.n_points = mi.desc.n_points;
    d.n_tris = mi.desc.n_tris;
    if (mi.on_device) {
      d.pts = mi.desc.xyz;
      d.tri = mi.desc.tri;
    } else {
      float* dp = (float*)(staging + soff);
      soff += align256(12 * mi.desc.n_points);
      int32_t* dt = (int32_t*)(staging + soff);
      soff += align256(12 * mi.desc.n_tris);
      HIP_TRY(hipMemcpyAsync(dp, mi.desc.xyz, 12 * mi.desc.n_points, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(dt, mi.desc.tri, 12 * mi.desc.n_tris, hipMemcpyHostToDevice, s));
      d.pts = dp;
      d.tri = dt;
    }
    for (int a = 0; a < 3; ++a) d.res[a] = mi.params.res[a];
    d.small_max = plan.tuning.class_small_max[size_class_of(d.res, d.n_tris)];
    d.tri_base = tri_base[i];
    d.pt_base = 0;
    tri_base[i + 1] = tri_base[i] + d.n_tris;
    max_points = std::max<uint64_t>(max_points, d.n_points);
  }
  const uint64_t T = tri_base[n];
  if (T >= (1ull << 32)) return SVO_E_RES_LIMIT;  // triangle indices are 32-bit on the device
  // per-class limits shrink with the pass's triangle count as the global one
  // does (below)
  for (MeshDev& d : md)
    if (d.small_max)
      d.small_max = (uint32_t)std::min<uint64_t>(d.small_max, std::max<uint64_t>(32, T / 128));

  // occupancy-bitmap words per mesh: one bit per Morton code below
  // morton(res - 1) + 1 (the reference's dense array size, SVO.cu:342-345),
  // each mesh's region rounded up to whole tiles of the level-1 pass
  std::vector<uint64_t> bm_off(n + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    const MeshDev& d = md[i];
    uint64_t g = 1;
    if (d.res[0] > 0 && d.res[1] > 0 && d.res[2] > 0)
      g = (uint64_t)morton3<uint64_t>((uint32_t)d.res[0] - 1, (uint32_t)d.res[1] - 1,
                                      (uint32_t)d.res[2] - 1) + 1;
    bm_off[i + 1] = bm_off[i] + bitmap_region_words(g);
  }
  const uint64_t bm_words = bm_off[n];

  // small control buffers: mesh descriptors, tri bases, bitmap offsets, bbox,
  // params, counters
  const size_t sz_md = align256(sizeof(MeshDev) * n), sz_tb = align256(8 * (n + 1)),
               sz_bo = align256(8 * (n + 1)), sz_acc = align256(32 * n),
               sz_par = align256(sizeof(MeshParams) * n), sz_cnt = align256(8 * 8);
  char* ctl = nullptr;
  SVO_TRY(W.get("ctl", sz_md + sz_tb + sz_bo + sz_acc + sz_par + sz_cnt, (void**)&ctl));
  MeshDev* d_md = (MeshDev*)ctl;
  uint64_t* d_tb = (uint64_t*)(ctl + sz_md);
  uint64_t* d_bo = (uint64_t*)(ctl + sz_md + sz_tb);
  uint32_t* d_acc = (uint32_t*)(ctl + sz_md + sz_tb + sz_bo);
  MeshParams* d_par = (MeshParams*)(ctl + sz_md + sz_tb + sz_bo + sz_acc);
  unsigned long long* d_cnt =
      (unsigned long long*)(ctl + sz_md + sz_tb + sz_bo + sz_acc + sz_par);
  // d_cnt: [0] box cells, [1] big chunks, [2] keys emitted, [3] big queued,
  // [4] bitmap bytes zeroed by the level-1 pass
  if (!capture) {
    SVO_TRY(L.ensure_pinned(sz_md + sz_tb + sz_bo + 64 + sizeof(MeshParams) * n));
    // stage the mesh descriptors, triangle bases and bitmap offsets through
    // the pinned buffer in one copy; the host does not touch the buffer again
    // before the next stream synchronization
    std::memcpy(L.pinned, md.data(), sizeof(MeshDev) * n);
    std::memcpy((char*)L.pinned + sz_md, tri_base.data(), 8 * (n + 1));
    std::memcpy((char*)L.pinned + sz_md + sz_tb, bm_off.data(), 8 * (n + 1));
    HIP_TRY(hipMemcpyAsync(d_md, L.pinned, sz_md + sz_tb + sz_bo, hipMemcpyHostToDevice, s));
  }
  launch_fill_u32((uint32_t*)d_cnt, 0u, 16, s);
  timer.mark("start");

  // ---- A0: bbox, params, classification
  launch_bbox(d_md, n, max_points, d_acc, s);
  launch_params(d_md, n, d_acc, d_par, s);
  // Triangles with at most small_max_cells box cells are voxelized one lane
  // each, larger ones by whole waves over 512-row items.  A pass with few
  // triangles has too few lanes to hide a long per-lane loop, so its limit
  // shrinks with the triangle count: T / 128 cells (32 below 4K triangles,
  // 512 at 64K, the default 1024 from 128K).
  Tuning tn = plan.tuning;
  tn.small_max_cells = (uint32_t)std::min<uint64_t>(tn.small_max_cells,
                                                    std::max<uint64_t>(32, T / 128));
  uint32_t* order = nullptr;
  SVO_TRY(W.get("order", order_bytes(T), (void**)&order));
  launch_classify(d_md, d_par, d_tb, n, T, tn, d_cnt, order,
                  (uint8_t*)order + align256(4 * T), s);
  timer.mark("classify");
  uint64_t cells = 0, n_big = 0;
  if (capture) {
    cells = rec->a0[0];
    n_big = rec->a0[1];
    launch_check_equal(d_cnt, rec->d_expect, 16, rec->d_flag, s);
  } else {
    HIP_TRY(hipMemcpyAsync(L.pinned, d_cnt, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    cells = ((uint64_t*)L.pinned)[0];
    n_big = ((uint64_t*)L.pinned)[1];
    if (rec) {
      rec->a0[0] = cells;
      rec->a0[1] = n_big;
    }
  }
  if (cells >= (1ull << 32) - 1) return SVO_E_OOM;  // key counts are 32-bit in the sort
  const uint64_t key_cap = std::max<uint64_t>(cells, 1);

  // ---- de-duplication path: occupancy bitmap (level 1 read straight from
  // it) while it is small enough, else a radix sort of the raw hits
  const uint64_t bm_bytes = 4 * bm_words;
  bool use_bm = bm_bytes <= plan.tuning.bitmap_max_bytes &&
                bm_bytes <= (uint64_t)plan.tuning.bitmap_bytes_per_cell * key_cap;
  if (plan.tuning.force_path == 1) use_bm = true;
  if (plan.tuning.force_path == 2) use_bm = false;
  // level-1 capacity: unique voxels <= min(candidate cells, bitmap bits).  A
  // pass of a plan sizes its levels for every grid cell and takes the bitmap
  // path (whose writes stay inside the grid), so that a replay over meshes
  // whose sizes changed writes nothing out of bounds before its checks stop it.
  if (rec) {
    if (bm_bytes > plan.tuning.bitmap_max_bytes) return SVO_E_RES_LIMIT;
    use_bm = true;
  }
  const uint64_t u_cap =
      use_bm ? std::max<uint64_t>(1, rec ? 32 * bm_words : std::min<uint64_t>(key_cap, 32 * bm_words))
             : key_cap;

  // ---- workspace for A1
  K *keys_a = nullptr, *keys_b = nullptr;
  uint32_t* bitmap = nullptr;
  BigItem* big = nullptr;
  void *sort_tmp = nullptr, *lvl_tmp = nullptr, *bm_tmp = nullptr;
  if (use_bm) {
    SVO_TRY(W.get("bitmap", bm_bytes, (void**)&bitmap));
    SVO_TRY(W.get("bm_tmp", bitmap_level_tmp_bytes(bm_words), &bm_tmp));
  } else {
    SVO_TRY(W.get("keys_a", sizeof(K) * key_cap, (void**)&keys_a));
    SVO_TRY(W.get("keys_b", sizeof(K) * key_cap, (void**)&keys_b));
  }
  SVO_TRY(W.get("big", sizeof(BigItem) * std::max<uint64_t>(n_big, 1), (void**)&big));
  // fixed grids sized from host-known capacities (small inputs get small grids)
  auto grid_for = [](uint64_t cap_elems, uint64_t per_block, uint32_t max_blocks) {
    const uint64_t b = (cap_elems + per_block - 1) / per_block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, max_blocks));
  };
  if (!use_bm) SVO_TRY(W.get("sort_tmp", radix_sort_tmp_bytes(key_cap), &sort_tmp));
  SVO_TRY(W.get("lvl_tmp", level_tmp_bytes(use_bm ? u_cap : key_cap), &lvl_tmp));
  // level capacities: cap_1 = unique voxels bound (box cells >= hits >=
  // unique voxels); cap_l = min(cap_{l-1}, number of level-l cells)
  std::vector<uint64_t> cap(NL + 2, 0);
  cap[1] = u_cap;
  for (uint32_t l = 2; l <= NL + 1; ++l) {
    uint64_t cellsl = 0;
    for (uint32_t i = 0; i < n; ++i) {
      uint64_t c = 1;
      for (int a = 0; a < 3; ++a) {
        const uint64_t r = (uint64_t)md[i].res[a];
        const uint64_t sh = l - 1;
        c *= sh >= 63 ? 1 : std::max<uint64_t>(1, (r + (1ull << sh) - 1) >> sh);
      }
      cellsl += c;
    }
    cap[l] = std::max<uint64_t>(1, std::min(cap[l - 1], cellsl));
  }
  std::vector<LevelArrays> lv(NL + 2);
  {
    size_t total = 0;
    // level 1 keeps only its keys: its group index per key is never read
    for (uint32_t l = 1; l <= NL + 1; ++l)
      total += align256(sizeof(K) * cap[l]) + (l >= 2 ? 2 * align256(4 * cap[l]) : 0);
    char* base = nullptr;
    SVO_TRY(W.get("levels", total, (void**)&base));
    for (uint32_t l = 1; l <= NL + 1; ++l) {
      lv[l].cap = cap[l];
      lv[l].elem = base;
      base += align256(sizeof(K) * cap[l]);
      if (l >= 2) {
        lv[l].elem_group = (uint32_t*)base;
        base += align256(4 * cap[l]);
        lv[l].first_child = (uint32_t*)base;
        base += align256(4 * cap[l]);
      } else {
        lv[l].elem_group = nullptr;
        lv[l].first_child = nullptr;
      }
    }
  }
  const size_t sz_counts = align256(8 * (NL + 2)), sz_spans = align256(sizeof(LevelSpan) * n * NL),
               sz_state = align256(sizeof(MeshTreeState) * n);
  char* lctl = nullptr;
  SVO_TRY(W.get("lctl", sz_counts + sz_spans + sz_state, (void**)&lctl));
  unsigned long long* d_counts = (unsigned long long*)lctl;
  LevelSpan* d_spans = (LevelSpan*)(lctl + sz_counts);
  MeshTreeState* d_state = (MeshTreeState*)(lctl + sz_counts + sz_spans);
  launch_fill_u32((uint32_t*)lctl, 0u, (sz_counts + sz_spans + sz_state) / 4, s);

  // ---- A1: voxelize, sort, levels
  timer.mark("setup_a1");
  // The level-1 pass zeroes the bitmap words it reads, so the lane's bitmap
  // stays clear between passes; only words never cleared since the buffer
  // was allocated are filled here (the stage is absent otherwise).  Until
  // the level-1 pass is launched the bitmap counts as dirty.
  size_t bm_clean_after = 0, bm_fill = 0;
  if (use_bm) {
    size_t& clean = W.clean_bytes("bitmap");
    if (clean < bm_bytes) {
      bm_fill = bm_bytes - clean;
      launch_fill_u32(bitmap + clean / 4, 0u, bm_fill / 4, s);
      timer.mark("bitmap_clear");
    }
    bm_clean_after = std::max(clean, (size_t)bm_bytes);
    clean = 0;
  }
  launch_voxelize<K>(d_md, d_par, d_tb, n, T, plan.cb, tn, keys_a, bitmap, d_bo,
                     d_cnt + 2, order, big, d_cnt + 3, n_big, s);
  timer.mark("voxelize");
  if (use_bm) {
    launch_level_first_bitmap<K>(bitmap, bm_words, d_bo, n, kl, lv[1], lv[2], d_counts, d_spans,
                                 NL, d_state, bm_tmp, lvl_tmp, d_cnt + 4, s);
    W.clean_bytes("bitmap") = bm_clean_after;
  } else {
    K* sorted = nullptr;
    launch_radix_sort<K>(keys_a, keys_b, d_cnt + 2, key_cap, plan.cb + plan.mb,
                         plan.tuning.sort_blocks, sort_tmp, &sorted, s);
    timer.mark("sort");
    launch_level_first<K>(sorted, d_cnt + 2, key_cap, kl, lv[1], lv[2], d_counts, d_spans, NL,
                          plan.tuning.scan_blocks, lvl_tmp, s);
    launch_level_first_finalize(d_spans, n, NL, d_state, s);
  }
  timer.mark("unique_level1");
  for (uint32_t l = 2; l <= NL; ++l)
    launch_level<K>((int)l, kl, lv[l], lv[l + 1], d_counts, d_spans, NL, d_state,
                    plan.tuning.scan_blocks, lvl_tmp, s);
  timer.mark("levels");

  // ---- readback of the level phase
  // counts, spans, states, params, then the three 8-byte hit counters on an
  // 8-byte boundary (76-byte MeshParams leave an odd count 4-byte aligned)
  const size_t off_spans = 8 * (NL + 2), off_state = off_spans + sizeof(LevelSpan) * n * NL,
               off_par = off_state + sizeof(MeshTreeState) * n,
               off_hits = (off_par + sizeof(MeshParams) * n + 7) & ~(size_t)7;
  char* hp = nullptr;
  if (capture) {
    if (rec->a1.size() != off_hits + 24) return SVO_E_INVALID;
    hp = rec->a1.data();
    const char* e1 = rec->d_expect + 16;
    launch_check_equal(d_counts, e1, 8 * (NL + 2), rec->d_flag, s);
    launch_check_equal(d_spans, e1 + off_spans, sizeof(LevelSpan) * n * NL, rec->d_flag, s);
    launch_check_equal(d_state, e1 + off_state, sizeof(MeshTreeState) * n, rec->d_flag, s);
    launch_check_equal(d_par, e1 + off_par, sizeof(MeshParams) * n, rec->d_flag, s);
    launch_check_equal(d_cnt + 2, e1 + off_hits, 24, rec->d_flag, s);
  } else {
    SVO_TRY(L.ensure_pinned(off_hits + 24));
    hp = (char*)L.pinned;
  }
  uint64_t* h_counts = (uint64_t*)hp;
  LevelSpan* h_spans = (LevelSpan*)(hp + off_spans);
  MeshTreeState* h_state = (MeshTreeState*)(hp + off_state);
  MeshParams* h_par = (MeshParams*)(hp + off_par);
  uint64_t* h_hits = (uint64_t*)(hp + off_hits);
  if (!capture) {
    HIP_TRY(hipMemcpyAsync(h_counts, d_counts, 8 * (NL + 2), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_spans, d_spans, sizeof(LevelSpan) * n * NL, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_state, d_state, sizeof(MeshTreeState) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_par, d_par, sizeof(MeshParams) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_hits, d_cnt + 2, 24, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (rec) rec->a1.assign(hp, hp + off_hits + 24);
  }
  const uint64_t n_hits = h_hits[0], bm_reset = h_hits[2];
  if (n_hits > key_cap) return SVO_E_HIP;  // cannot happen: hits <= box cells
  // the pinned area is reused below: keep host copies of the readback
  const std::vector<uint64_t> v_counts(h_counts, h_counts + NL + 2);
  const std::vector<LevelSpan> v_spans(h_spans, h_spans + (size_t)n * NL);
  const std::vector<MeshTreeState> v_state(h_state, h_state + n);
  const std::vector<MeshParams> v_par(h_par, h_par + n);
  h_counts = const_cast<uint64_t*>(v_counts.data());
  h_spans = const_cast<LevelSpan*>(v_spans.data());
  h_state = const_cast<MeshTreeState*>(v_state.data());
  h_par = const_cast<MeshParams*>(v_par.data());

  // ---- per-mesh node layout
  std::vector<MeshNodeLayout> lay(n);
  std::vector<uint64_t> node_base(n + 1, 0);
  std::vector<int> mstat(n, SVO_OK);
  uint64_t N = 0;
  for (uint32_t i = 0; i < n; ++i) {
    MeshNodeLayout& Lm = lay[i];
    std::memset(&Lm, 0, sizeof(Lm));
    const MeshParams& P = h_par[i];
    const uint32_t d = h_state[i].depth;
    if (P.status != 0) mstat[i] = SVO_E_INVALID;
    else if (d == 0) mstat[i] = SVO_E_EMPTY;
    Lm.node_base = N;
    if (mstat[i] == SVO_OK) {
      Lm.depth = d;
      uint64_t run = 0;
      for (uint32_t l = 1; l < d; ++l) {
        const LevelSpan& sp = h_spans[(uint64_t)i * NL + (l - 1)];
        Lm.level_off[l - 1] = run;
        run += 8ull * (sp.group_end - sp.group_first);
      }
      Lm.level_off[d - 1] = run;
      Lm.n_nodes = run + 1;
      float w = P.unit[0];  // unitNodeWidth = unitVoxelSize.x(), doubled per level (SVO.cu:483)
      for (uint32_t l = 0; l < 24; ++l) {
        Lm.width[l] = w;
        w = (float)((double)w * 2.0);
      }
      if (Lm.n_nodes >= (1ull << 32)) mstat[i] = SVO_E_RES_LIMIT;  // uint32 node links
    }
    if (mstat[i] != SVO_OK) Lm.n_nodes = 0;
    Lm.key_out_bits = plan.code_bytes * 8;
    N += Lm.n_nodes;
    node_base[i + 1] = N;
  }
  const uint64_t U1 = h_counts[0];
  const int sem = (int)in[plan.idx[0]].params.semantics;
  const bool want_attrs = !(in[plan.idx[0]].params.flags & SVO_BUILD_NO_VERTEX_EDGE);

  // ---- output block
  const int cbytes = plan.code_bytes;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const size_t o_code = take(cbytes * N), o_leaf = take(N), o_org = take(12 * N),
               o_w = take(4 * N), o_par = take(4 * N), o_ch = take(32 * N),
               o_nb = take(108 * N), o_vx = take(cbytes * U1);
  const size_t o_v = take(want_attrs ? 16 * 8 * N : 0);
  const size_t o_e = take(want_attrs ? 28 * 12 * N : 0);
  const size_t o_f = take(want_attrs && sem == SVO_SEM_INTENDED ? 28 * 6 * N : 0);
  std::shared_ptr<BatchOutput> out;
  if (capture) {
    // every replay writes the recorded run's block
    if (!rec->out || rec->out->block->bytes != off) return SVO_E_INVALID;
    out = rec->out;
  } else {
    out = std::make_shared<BatchOutput>();
    out->block = std::make_shared<DeviceBlock>();
    void* p = nullptr;
    SVO_TRY(C.dev_malloc(&p, off));
    out->block->ctx = C.shared_from_this();
    out->block->ptr = p;
    out->block->bytes = off;
    if (rec) rec->out = out;
  }
  // From here on kernels of this pass write the output block.  On an error
  // return, wait for both streams before `out` (declared above, so destroyed
  // after this guard) gives the block back to the context's cache, where
  // another lane could take it while this pass's kernels still write it.
  struct DrainOnError {
    Lane& L;
    bool armed = true;
    ~DrainOnError() {
      if (!armed) return;
      hipStreamSynchronize(L.stream);
      hipStreamSynchronize(L.side);
      hipGetLastError();
    }
  } drain{L};
  char* ob = (char*)out->block->ptr;
  NodeArrays& na = out->nodes;
  na.code = ob + o_code;
  na.is_leaf = (uint8_t*)(ob + o_leaf);
  na.origin = (float*)(ob + o_org);
  na.width = (float*)(ob + o_w);
  na.parent = (uint32_t*)(ob + o_par);
  na.childs = (uint32_t*)(ob + o_ch);
  na.neighbors = (uint32_t*)(ob + o_nb);
  na.code_bytes = cbytes;
  na.fill_neighbors = 0;
  na.skip = capture ? rec->d_flag : nullptr;
  out->voxels = ob + o_vx;
  out->verts = (svo_vertex*)(ob + o_v);
  out->edges = (svo_edge*)(ob + o_e);
  out->faces = (svo_edge*)(ob + o_f);
  out->code_bytes = cbytes;

  // ---- B: node records, neighbors, vertex/edge/face records
  char* bctl = nullptr;
  const size_t sz_lay = align256(sizeof(MeshNodeLayout) * n), sz_nb = align256(8 * (n + 1)),
               sz_tot = align256(8 * (4 + 3 * n));
  SVO_TRY(W.get("bctl", sz_lay + sz_nb + sz_tot, (void**)&bctl));
  MeshNodeLayout* d_lay = (MeshNodeLayout*)bctl;
  uint64_t* d_nbase = (uint64_t*)(bctl + sz_lay);
  unsigned long long* d_tot = (unsigned long long*)(bctl + sz_lay + sz_nb);
  if (!capture) {
    SVO_TRY(L.ensure_pinned(sz_lay + sz_nb + sz_tot));
    std::memcpy(L.pinned, lay.data(), sizeof(MeshNodeLayout) * n);
    std::memcpy((char*)L.pinned + sz_lay, node_base.data(), 8 * (n + 1));
    HIP_TRY(hipMemcpyAsync(bctl, L.pinned, sz_lay + sz_nb, hipMemcpyHostToDevice, s));
  }
  launch_fill_u32((uint32_t*)d_tot, 0u, sz_tot / 4, s);
  timer.mark("setup_b");
  // When this pass runs alone on the context, the neighbor fill (a pure
  // write stream) runs on the side stream beside the node records (bound by
  // their loads' latency); the roots' neighbor slot is written after the
  // join.  Concurrent passes already overlap each other, and their extra
  // streams would outnumber the hardware queues: there the node and root
  // kernels write the neighbor slots themselves (wave-coalesced stores
  // beside their record writes).
  bool side_fill = C.active_passes.load() == 1 && !capture;
  if (side_fill && (hipEventRecord(L.fork, s) != hipSuccess ||
                    hipStreamWaitEvent(L.side, L.fork, 0) != hipSuccess)) {
    hipGetLastError();
    side_fill = false;  // fill in order instead
  }
  if (side_fill) {
    const size_t e0 = timer.side_mark(L.side);
    launch_fill_u32(na.neighbors, NONE, 27 * N, L.side);
    const size_t e1 = timer.side_mark(L.side);
    timer.side_range("neighbor_fill", e0, e1);
    if (hipEventRecord(L.join, L.side) != hipSuccess) {
      hipGetLastError();
      hipStreamSynchronize(L.side);  // no join event: wait for the fill here
      side_fill = false;
    }
  } else {
    na.fill_neighbors = 1;
  }
  for (uint32_t l = 1; l < NL; ++l)
    launch_emit_level_nodes<K>((int)l, kl, lv[l].elem, h_counts[l - 1], lv[l + 1].elem,
                               lv[l + 1].elem_group, lv[l + 1].first_child, h_counts[l], d_spans,
                               NL, d_state, d_lay, d_par, na, s);
  if (side_fill && hipStreamWaitEvent(s, L.join, 0) != hipSuccess) {
    hipStreamSynchronize(L.side);  // the fill writes the output block: let it end first
    return SVO_E_HIP;
  }
  {
    std::vector<void*> lp(NL + 1, nullptr);
    for (uint32_t l = 1; l <= NL; ++l) lp[l - 1] = lv[l].elem;
    launch_emit_roots<K>(kl, lp.data(), n, d_spans, NL, d_state, d_lay, d_par, sem, na, s);
  }
  launch_copy_voxels<K>(kl, lv[1].elem, U1, out->voxels, cbytes, s);
  timer.mark("nodes");
  if (sem == SVO_SEM_INTENDED)
    for (uint32_t l = NL - 1; l >= 1; --l)
      launch_neighbors_level<K>((int)l, kl, lv[l + 1].elem, lv[l + 1].elem_group, h_counts[l],
                                d_spans, NL, d_state, d_lay, na, s);
  timer.mark("neighbors");
  void* attr_tmp = nullptr;
  if (want_attrs && N) {
    AttrOut ao{out->verts, out->edges, out->faces, d_tot};
    if (sem == SVO_SEM_REFERENCE && no_adjacent_duplicates(lay, h_par, mstat)) {
      timer.mark("attr_count");
      launch_ref_attrs(na, N, d_nbase, n, ao, s);
    } else {
      SVO_TRY(W.get("attr_tmp", attr_tmp_bytes(plan.tuning.scan_blocks), &attr_tmp));
      const uint32_t ab = grid_for(N, 2048, plan.tuning.scan_blocks);
      launch_attrs_count(sem, na, N, d_nbase, n, ab, attr_tmp, ao, s);
      timer.mark("attr_count");
      launch_attrs_write(sem, na, N, d_nbase, n, ab, attr_tmp, ao, s);
    }
  }
  timer.mark("attr_write");
  std::vector<uint64_t> tot;
  if (capture) {
    if (rec->tot.size() != 4 + 3 * (size_t)n) return SVO_E_INVALID;
    tot = rec->tot;
    launch_check_equal(d_tot, rec->d_expect + 16 + rec->a1.size(), 8 * (4 + 3 * n), rec->d_flag,
                       s);
  } else {
    uint64_t* h_tot = (uint64_t*)L.pinned;
    HIP_TRY(hipMemcpyAsync(h_tot, d_tot, 8 * (4 + 3 * n), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    tot.assign(h_tot, h_tot + 4 + 3 * n);
    if (rec) rec->tot = tot;
  }

  // ---- trees
  if (!capture) {
    svo_pass_stats ps{};
    ps.n_meshes = n;
    ps.n_tris = T;
    ps.box_cells = cells;
    ps.n_hits = n_hits;
    ps.n_voxels = U1;
    for (uint32_t l = 0; l < NL; ++l) ps.level_elems += h_counts[l];
    ps.n_nodes = N;
    ps.n_vertices = tot[0];
    ps.n_edges = tot[1];
    ps.n_faces = tot[2];
    ps.out_bytes = out->block->bytes;
    ps.key_bytes = sizeof(K);
    ps.sort_bits = (uint32_t)(plan.cb + plan.mb);
    ps.sort_passes = use_bm ? 0 : (ps.sort_bits + 7) / 8;
    ps.bitmap_bytes = use_bm ? bm_bytes : 0;
    ps.n_groups1 = h_counts[1];
    ps.bitmap_fill_bytes = bm_fill;
    ps.bitmap_reset_bytes = bm_reset;
    ps.n_levels = NL;
    timer.publish(*plan.timing, ps);
  }
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t src = plan.idx[i];
    statuses[src] = mstat[i];
    if (mstat[i] != SVO_OK) continue;
    svo_tree* t = capture ? &rec->templ[src] : new svo_tree();
    t->out = out;
    t->ctx = C.shared_from_this();
    t->node_base = node_base[i];
    t->semantics = sem;
    t->key_bits_out = in[src].params.key_bits ? (int)in[src].params.key_bits
                                               : (plan.code_bytes == 8 ? 64 : 32);
    int kbv = 32;
    validate_params(&in[src].desc, &in[src].params, &kbv);
    t->key_bits_out = kbv;
    svo_tree_info& I = t->info;
    I.depth = (int32_t)lay[i].depth;
    I.key_bits = kbv;
    I.semantics = sem;
    const MeshParams& P = h_par[i];
    for (int a = 0; a < 3; ++a) {
      I.res[a] = P.res[a];
      I.bbox_min[a] = P.bmin[a];
      I.bbox_max[a] = P.bmax[a];
      I.bbox_width[a] = P.width[a];
      I.unit[a] = P.unit[a];
    }
    const LevelSpan& s1 = h_spans[(uint64_t)i * NL];
    t->vox_base = s1.elem_first;
    I.n_voxels = s1.elem_end - s1.elem_first;
    I.n_nodes = lay[i].n_nodes;
    for (uint32_t l = 0; l < lay[i].depth && l < 32; ++l)
      I.level_nodes[l] = (l + 1 < lay[i].depth) ? lay[i].level_off[l + 1] - lay[i].level_off[l] : 1;
    if (want_attrs) {
      // records of mesh i run from its first node's offsets to the next mesh's
      uint64_t nv = tot[0], ne = tot[1], nf = tot[2];
      for (uint32_t j = i + 1; j < n; ++j)
        if (mstat[j] == SVO_OK && lay[j].n_nodes) {
          nv = tot[4 + 3 * j];
          ne = tot[4 + 3 * j + 1];
          nf = tot[4 + 3 * j + 2];
          break;
        }
      t->v_base = tot[4 + 3 * i];
      t->e_base = tot[4 + 3 * i + 1];
      t->f_base = tot[4 + 3 * i + 2];
      I.n_vertices = nv - t->v_base;
      I.n_edges = ne - t->e_base;
      I.n_faces = nf - t->f_base;
    }
    I.n_hits = n == 1 ? n_hits : 0;
    I.device_bytes = out->block->bytes;
    if (!capture) trees[src] = t;
  }
  if (capture) {
    for (uint32_t i = 0; i < n; ++i) rec->templ_status[plan.idx[i]] = mstat[i];
    drain.armed = false;  // nothing ran: the graph is being recorded
    return SVO_OK;
  }
  drain.armed = false;  // both streams were synchronized above
  return SVO_OK;
}

}  // namespace

namespace {

// The pass plan of meshes `sel` of `in`: key layout, levels, tuning snapshot.
// SVO_E_RES_LIMIT when their composite keys do not fit 64 bits.
int make_plan(Context& C, const std::vector<MeshInput>& in, const std::vector<uint32_t>& sel,
              Plan& plan) {
  bool any64 = false;
  for (uint32_t i : sel) {
    plan.idx.push_back(i);
    any64 |= in[i].params.key_bits == 64 ||
             (in[i].params.key_bits == 0 &&
              std::max({in[i].params.res[0], in[i].params.res[1], in[i].params.res[2]}) > 1024);
    for (int a = 0; a < 3; ++a) plan.bits = std::max(plan.bits, bits_for(in[i].params.res[a]));
  }
  plan.cb = 3 * plan.bits;
  plan.mb = bits_for((int32_t)plan.idx.size());
  plan.n_levels = plan.bits + 1;
  if (plan.cb + plan.mb > 64) return SVO_E_RES_LIMIT;
  plan.k64 = plan.cb + plan.mb > 32;
  plan.code_bytes = any64 ? 8 : 4;
  std::lock_guard<std::mutex> lk(C.mu);
  plan.tuning = C.tuning;
  return SVO_OK;
}

int run_plan_typed(Context& C, Lane& L, const std::vector<MeshInput>& in, const Plan& plan,
                   std::vector<svo_tree*>& trees, std::vector<int>& statuses) {
  return plan.k64 ? run_typed<uint64_t>(C, L, in, plan, trees, statuses)
                  : run_typed<uint32_t>(C, L, in, plan, trees, statuses);
}

// One pipeline pass over the valid meshes `sel` of `in` on one lane.
int run_pass(Context& C, const std::vector<MeshInput>& in, const std::vector<uint32_t>& sel,
             std::vector<svo_tree*>& trees, std::vector<int>& statuses, CallTiming& ct,
             int lane_pref = -1) {
  struct ActivePass {
    Context& c;
    explicit ActivePass(Context& x) : c(x) { c.active_passes.fetch_add(1); }
    ~ActivePass() { c.active_passes.fetch_sub(1); }
  } active(C);
  if (sel.empty()) return SVO_OK;
  Plan plan;
  plan.timing = &ct;
  if (make_plan(C, in, sel, plan) != SVO_OK) {
    for (uint32_t i : plan.idx) statuses[i] = SVO_E_RES_LIMIT;
    return SVO_E_RES_LIMIT;
  }
  Lane* L = C.acquire_lane(lane_pref);
  if (!L) return SVO_E_HIP;
  hipSetDevice(C.device);
  int st = run_plan_typed(C, *L, in, plan, trees, statuses);
  if (st != SVO_OK) {
    // the lane's workspaces are reused by its next pass: let this one end
    hipStreamSynchronize(L->stream);
    hipStreamSynchronize(L->side);
    hipGetLastError();
    for (uint32_t i : plan.idx) {
      delete trees[i];
      trees[i] = nullptr;
      statuses[i] = st;
    }
  }
  C.release_lane(L);
  return st;
}

// Meshes `valid` split into the sub-batches of concurrent passes: large
// batches are split so that the passes' host round trips and kernel tails
// overlap.  Meshes go to the lightest sub-batch, heaviest first; the weight
// estimates triangles plus surface voxels (~ res^2).
std::vector<std::vector<uint32_t>> split_batch(Context& C, const std::vector<MeshInput>& in,
                                               const std::vector<uint32_t>& valid) {
  uint32_t groups = 1;
  {
    Tuning t;
    {
      std::lock_guard<std::mutex> lk(C.mu);
      t = C.tuning;
    }
    if (t.batch_lanes > 1 && valid.size() >= 2 * (size_t)t.batch_min_meshes)
      groups = std::min<uint32_t>(t.batch_lanes, (uint32_t)(valid.size() / t.batch_min_meshes));
  }
  if (groups <= 1) return {valid};
  std::vector<std::pair<double, uint32_t>> w;
  for (uint32_t i : valid) {
    const double r = std::max({in[i].params.res[0], in[i].params.res[1], in[i].params.res[2]});
    w.push_back({3.0 * (double)in[i].desc.n_tris + 10.0 * r * r, i});
  }
  std::sort(w.begin(), w.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<std::vector<uint32_t>> sel(groups);
  std::vector<double> load(groups, 0.0);
  for (const auto& e : w) {
    const uint32_t g = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
    sel[g].push_back(e.second);
    load[g] += e.first;
  }
  for (auto& v : sel) std::sort(v.begin(), v.end());
  return sel;
}

// Meshes `valid` (one semantics and flag set) as one pass, or as concurrent
// sub-batch passes on separate lanes when there are enough of them.
int run_set(Context& C, const std::vector<MeshInput>& in, const std::vector<uint32_t>& valid,
            std::vector<svo_tree*>& trees, std::vector<int>& statuses, CallTiming& ct) {
  const std::vector<std::vector<uint32_t>> sel = split_batch(C, in, valid);
  const uint32_t groups = (uint32_t)sel.size();
  if (groups <= 1) return run_pass(C, in, valid, trees, statuses, ct);
  std::vector<int> rc(groups, SVO_OK);
  std::vector<std::thread> th;
  for (uint32_t g = 1; g < groups; ++g)
    th.emplace_back([&, g] { rc[g] = run_pass(C, in, sel[g], trees, statuses, ct, (int)g); });
  rc[0] = run_pass(C, in, sel[0], trees, statuses, ct, 0);
  for (auto& t : th) t.join();
  for (int r : rc)
    if (r != SVO_OK) return r;
  return SVO_OK;
}

}  // namespace

// ------------------------------------------------------------------ batch plans
// A batch recorded once and replayed as one HIP graph: an eager run of each
// sub-batch pass on a lane of the plan's own records the host values of its
// three readbacks (PassRecord) and sizes the lane's workspaces; the same
// passes are then captured -- the sub-batches as parallel branches forked
// from one stream -- with the readbacks replaced by device-side comparisons
// against the recorded values.  A replay is one graph launch and one stream
// synchronization; it recomputes every output from the (device-resident)
// inputs into the recorded output blocks, and a set flag (a count, span or
// total that differs from the recorded run) makes svo_plan_run fall back to
// an eager build.  Replaces the per-level host synchronizations of the
// reference's level loop (SVO.cu:383-495) by one per batch.
BatchPlan::~BatchPlan() {
  Context& C = *ctx;
  hipSetDevice(C.device);
  for (auto& l : lanes)
    if (l->stream) hipStreamSynchronize(l->stream);
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  if (fork) hipEventDestroy(fork);
  for (hipEvent_t e : joins)
    if (e) hipEventDestroy(e);
  for (auto& e : expect) C.dev_free(e.first, e.second);
  if (d_flag) C.dev_free(d_flag, 256);
  if (h_flag) hipHostFree(h_flag);
  recs.clear();  // output blocks go back to the context
  for (auto& l : lanes) {
    l->ws.release_all();
    for (auto ev : l->events) hipEventDestroy(ev);
    if (l->fork) hipEventDestroy(l->fork);
    if (l->join) hipEventDestroy(l->join);
    if (l->pinned) hipHostFree(l->pinned);
    if (l->stream) hipStreamDestroy(l->stream);
    if (l->side) hipStreamDestroy(l->side);
  }
  for (auto& o : owned) C.dev_free(o.first, o.second);
}

int plan_create(Context& C, const std::vector<MeshInput>& user, std::unique_ptr<BatchPlan>& out) {
  auto bp = std::make_unique<BatchPlan>();
  bp->ctx = C.shared_from_this();
  bp->in = user;
  HIP_TRY(hipSetDevice(C.device));
  const uint32_t n = (uint32_t)user.size();
  if (n == 0) return SVO_E_INVALID;
  for (uint32_t i = 0; i < n; ++i) {
    int kb = 32;
    SVO_TRY(validate_params(&user[i].desc, &user[i].params, &kb));
    if (user[i].params.semantics != user[0].params.semantics ||
        user[i].params.flags != user[0].params.flags)
      return SVO_E_INVALID;  // one semantics and flag set per plan
  }
  // lanes of the plan's own: nothing else may touch the workspaces the graph
  // points into
  std::vector<uint32_t> all(n);
  for (uint32_t i = 0; i < n; ++i) all[i] = i;
  const auto sel = split_batch(C, bp->in, all);
  if (sel.size() > 64) return SVO_E_INVALID;
  for (size_t g = 0; g < sel.size(); ++g) {
    auto l = std::make_unique<Lane>();
    if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&l->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&l->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&l->join, hipEventDisableTiming) != hipSuccess)
      return SVO_E_HIP;
    l->ws.ctx = &C;
    bp->lanes.push_back(std::move(l));
  }
  hipStream_t s0 = bp->lanes[0]->stream;
  // host meshes become device copies owned by the plan
  for (uint32_t i = 0; i < n; ++i) {
    MeshInput& mi = bp->in[i];
    if (mi.on_device) continue;
    void* p = nullptr;
    void* t = nullptr;
    SVO_TRY(C.dev_malloc(&p, 12 * mi.desc.n_points));
    bp->owned.push_back({p, 12 * mi.desc.n_points});
    SVO_TRY(C.dev_malloc(&t, 12 * mi.desc.n_tris));
    bp->owned.push_back({t, 12 * mi.desc.n_tris});
    HIP_TRY(hipMemcpyAsync(p, mi.desc.xyz, 12 * mi.desc.n_points, hipMemcpyHostToDevice, s0));
    HIP_TRY(hipMemcpyAsync(t, mi.desc.tri, 12 * mi.desc.n_tris, hipMemcpyHostToDevice, s0));
    HIP_TRY(hipStreamSynchronize(s0));
    mi.desc.xyz = (const float*)p;
    mi.desc.tri = (const int32_t*)t;
    mi.on_device = true;
  }
  // record: one eager run of every pass
  CallTiming ct;
  for (size_t g = 0; g < sel.size(); ++g) {
    Plan pl;
    pl.timing = &ct;
    SVO_TRY(make_plan(C, bp->in, sel[g], pl));
    auto rec = std::make_unique<PassRecord>();
    pl.rec = rec.get();
    std::vector<svo_tree*> trees(n, nullptr);
    std::vector<int> st(n, SVO_OK);
    const int rc = run_plan_typed(C, *bp->lanes[g], bp->in, pl, trees, st);
    for (svo_tree* t : trees) delete t;  // the record keeps the output block
    if (rc != SVO_OK) return rc;
    bp->plans.push_back(pl);
    bp->recs.push_back(std::move(rec));
  }
  // expected readbacks on the device, [a0 | a1 | tot] per pass
  for (auto& rec : bp->recs) {
    std::vector<char> img(16 + rec->a1.size() + 8 * rec->tot.size());
    std::memcpy(img.data(), rec->a0, 16);
    std::memcpy(img.data() + 16, rec->a1.data(), rec->a1.size());
    std::memcpy(img.data() + 16 + rec->a1.size(), rec->tot.data(), 8 * rec->tot.size());
    void* d = nullptr;
    SVO_TRY(C.dev_malloc(&d, img.size()));
    bp->expect.push_back({d, img.size()});
    HIP_TRY(hipMemcpyAsync(d, img.data(), img.size(), hipMemcpyHostToDevice, s0));
    HIP_TRY(hipStreamSynchronize(s0));
    rec->d_expect = (const char*)d;
  }
  SVO_TRY(C.dev_malloc((void**)&bp->d_flag, 256));
  HIP_TRY(hipHostMalloc((void**)&bp->h_flag, 256));
  HIP_TRY(hipEventCreateWithFlags(&bp->fork, hipEventDisableTiming));
  bp->joins.assign(sel.size(), nullptr);
  for (size_t g = 1; g < sel.size(); ++g)
    HIP_TRY(hipEventCreateWithFlags(&bp->joins[g], hipEventDisableTiming));
  for (size_t g = 0; g < sel.size(); ++g) {
    bp->lanes[g]->ws.frozen = true;
    PassRecord& r = *bp->recs[g];
    r.mode = PassRecord::CAPTURE;
    // a flag per pass: set only by the pass's own checks, which its later
    // kernels follow in stream order
    r.d_flag = bp->d_flag + g;
    r.templ.assign(n, svo_tree());
    r.templ_status.assign(n, SVO_E_INVALID);
  }
  bp->sel = sel;
  // capture: the passes as branches forked from lane 0's stream
  HIP_TRY(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  int rc = SVO_OK;
  launch_fill_u32(bp->d_flag, 0u, 64, s0);
  if (hipEventRecord(bp->fork, s0) != hipSuccess) rc = SVO_E_HIP;
  for (size_t g = 1; g < sel.size() && rc == SVO_OK; ++g)
    if (hipStreamWaitEvent(bp->lanes[g]->stream, bp->fork, 0) != hipSuccess) rc = SVO_E_HIP;
  for (size_t g = 0; g < sel.size() && rc == SVO_OK; ++g) {
    std::vector<svo_tree*> trees(n, nullptr);
    std::vector<int> st(n, SVO_OK);
    rc = run_plan_typed(C, *bp->lanes[g], bp->in, bp->plans[g], trees, st);
  }
  for (size_t g = 1; g < sel.size() && rc == SVO_OK; ++g)
    if (hipEventRecord(bp->joins[g], bp->lanes[g]->stream) != hipSuccess ||
        hipStreamWaitEvent(s0, bp->joins[g], 0) != hipSuccess)
      rc = SVO_E_HIP;
  if (rc == SVO_OK &&
      hipMemcpyAsync(bp->h_flag, bp->d_flag, 4 * sel.size(), hipMemcpyDeviceToHost, s0) !=
          hipSuccess)
    rc = SVO_E_HIP;
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s0, &graph);
  bp->graph = graph;
  if (rc != SVO_OK) {
    hipGetLastError();
    return rc;
  }
  if (ec != hipSuccess || !graph) return SVO_E_HIP;
  HIP_TRY(hipGraphInstantiate(&bp->exec, graph, nullptr, nullptr, 0));
  out = std::move(bp);
  return SVO_OK;
}

int plan_run(BatchPlan& bp, std::vector<svo_tree*>& trees, std::vector<int>& statuses,
             bool& replayed) {
  std::lock_guard<std::mutex> lk(bp.mu);
  Context& C = *bp.ctx;
  const uint32_t n = (uint32_t)bp.in.size();
  HIP_TRY(hipSetDevice(C.device));
  hipStream_t s0 = bp.lanes[0]->stream;
  const size_t np = bp.sel.size();
  for (size_t g = 0; g < np; ++g) bp.h_flag[g] = 0xFFFFFFFFu;  // overwritten by the graph
  HIP_TRY(hipGraphLaunch(bp.exec, s0));
  HIP_TRY(hipStreamSynchronize(s0));
  uint32_t any = 0;
  for (size_t g = 0; g < np; ++g) any |= bp.h_flag[g];
  if (any != 0) {
    // something the graph was recorded with changed: build eagerly
    replayed = false;
    return run_batch(C, bp.in, trees, statuses);
  }
  replayed = true;
  trees.assign(n, nullptr);
  statuses.assign(n, SVO_OK);
  for (size_t g = 0; g < bp.sel.size(); ++g) {
    const PassRecord& r = *bp.recs[g];
    for (uint32_t i : bp.sel[g]) {
      statuses[i] = r.templ_status[i];
      if (r.templ_status[i] == SVO_OK) trees[i] = new svo_tree(r.templ[i]);
    }
  }
  return SVO_OK;
}

int run_batch(Context& C, const std::vector<MeshInput>& in, std::vector<svo_tree*>& trees,
              std::vector<int>& statuses) {
  trees.assign(in.size(), nullptr);
  statuses.assign(in.size(), SVO_OK);
  // A pass computes one semantics and one flag set: the valid meshes are
  // grouped by them, in order of first appearance, one set of passes each.
  std::vector<std::vector<uint32_t>> sets;
  for (uint32_t i = 0; i < in.size(); ++i) {
    int kb = 32;
    const int st = validate_params(&in[i].desc, &in[i].params, &kb);
    statuses[i] = st;
    if (st != SVO_OK) continue;
    size_t k = 0;
    while (k < sets.size() && (in[sets[k][0]].params.semantics != in[i].params.semantics ||
                               in[sets[k][0]].params.flags != in[i].params.flags))
      ++k;
    if (k == sets.size()) sets.emplace_back();
    sets[k].push_back(i);
  }
  if (sets.empty()) return SVO_OK;
  CallTiming ct;
  if (C.timing) {
    hipSetDevice(C.device);
    if (StageTimer::create_event(&ct.ref) == hipSuccess) {
      // on the null stream, which the lanes' non-blocking streams do not
      // wait for and nothing else uses: it completes when the call begins
      if (hipEventRecord(ct.ref, nullptr) != hipSuccess) {
        hipEventDestroy(ct.ref);
        ct.ref = nullptr;
      }
    } else {
      ct.ref = nullptr;
    }
  }
  struct RefGuard {
    CallTiming& ct;
    Context& C;
    ~RefGuard() {
      ct.publish(C);
      if (ct.ref) hipEventDestroy(ct.ref);
    }
  } ref_guard{ct, C};
  int rc = SVO_OK;
  for (const auto& set : sets) {
    const int r = run_set(C, in, set, trees, statuses, ct);
    if (rc == SVO_OK) rc = r;
  }
  return rc;
}

}  // namespace svo
