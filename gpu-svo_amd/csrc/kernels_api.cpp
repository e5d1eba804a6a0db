// kernels_api.cpp -- include/svo/svo_kernels.h: single kernels on host buffers.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "svo/svo_kernels.h"
#include "svo_host.h"

using namespace svo;

namespace {

struct DevBufs {
  std::vector<void*> ptrs;
  ~DevBufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  void* get(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 256) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return p;
  }
};

// Fixed device scratch of svo_k_radix_sort_noalloc: the key count the sort
// kernels read, the ping-pong key buffer and the digit counts of up to
// NOALLOC_SORT_MAX_KEYS keys of 8 bytes.  Each device holds its own copy for
// the life of the process.  In host code the array's name is not a device
// address: kernels get the per-device address that hipGetSymbolAddress
// returned for it (noalloc_scratch below).
constexpr uint64_t NOALLOC_SORT_MAX_KEYS = 1ull << 22;
constexpr size_t NOALLOC_SCRATCH_BYTES = 34ull << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

__global__ void k_set_count(unsigned long long* c, unsigned long long v) {
  if (threadIdx.x == 0) *c = v;
}

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device; resolved on its first call

// Device address of `scratch` on `dev` (the current device).
int noalloc_scratch(int dev, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.empty()) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SVO_E_HIP;
    g_scratch_ptr.assign((size_t)n, nullptr);
  }
  if (dev < 0 || (size_t)dev >= g_scratch_ptr.size()) return SVO_E_INVALID;
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return SVO_E_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SVO_OK;
}

}  // namespace

namespace {
template <typename K>
int octree_level(const K* sorted, uint64_t n, int code_bits, K* unique, uint64_t* n_unique,
                 K* parents, uint32_t* first_child, uint64_t* n_parents) {
  // sorted, and every code below 2^code_bits (the span table has one row)
  const K lim = (K)1 << code_bits;
  for (uint64_t i = 0; i < n; ++i)
    if (sorted[i] >= lim || (i && sorted[i] < sorted[i - 1])) return SVO_E_INVALID;
  DevBufs B;
  K* in = (K*)B.get(sizeof(K) * n);
  K* e1 = (K*)B.get(sizeof(K) * n);
  K* e2 = (K*)B.get(sizeof(K) * n);
  uint32_t* fc = (uint32_t*)B.get(4 * n);
  unsigned long long* cnt = (unsigned long long*)B.get(8 * 4);  // [n, U, G, -]
  LevelSpan* spans = (LevelSpan*)B.get(sizeof(LevelSpan));
  void* tmp = B.get(level_tmp_bytes(n));
  if (!in || !e1 || !e2 || !fc || !cnt || !spans || !tmp) return SVO_E_OOM;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return SVO_E_HIP;
  int st = SVO_E_HIP;
  const unsigned long long hn[4] = {n, 0, 0, 0};
  unsigned long long hc[4] = {0, 0, 0, 0};
  if (hipMemcpyAsync(in, sorted, sizeof(K) * n, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(cnt, hn, sizeof hn, hipMemcpyHostToDevice, s) == hipSuccess) {
    const KeyLayout<K> kl{code_bits, (K)(lim - 1)};
    const LevelArrays l1{e1, nullptr, nullptr, n}, l2{e2, nullptr, fc, n};
    launch_level_first<K>(in, cnt, n, kl, l1, l2, cnt + 1, spans, 1, Tuning{}.scan_blocks, tmp, s);
    if (hipMemcpyAsync(hc, cnt, sizeof hc, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess && hipGetLastError() == hipSuccess &&
        hc[1] <= n && hc[2] <= n &&
        hipMemcpyAsync(unique, e1, sizeof(K) * hc[1], hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipMemcpyAsync(parents, e2, sizeof(K) * hc[2], hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipMemcpyAsync(first_child, fc, 4 * hc[2], hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess) {
      *n_unique = hc[1];
      *n_parents = hc[2];
      st = SVO_OK;
    }
  }
  (void)hipStreamDestroy(s);
  return st;
}
}  // namespace

extern "C" {

int svo_k_radix_sort(int device, void* keys, uint64_t n, int key_bytes, int bits) {
  if ((key_bytes != 4 && key_bytes != 8) || bits < 0 || bits > 8 * key_bytes ||
      (n && !keys) || n >= (1ull << 32))
    return SVO_E_INVALID;
  if (n == 0) return SVO_OK;
  if (hipSetDevice(device) != hipSuccess) return SVO_E_HIP;
  DevBufs B;
  const uint32_t blocks = Tuning{}.sort_blocks;
  void* a = B.get(key_bytes * n);
  void* b = B.get(key_bytes * n);
  void* tmp = B.get(radix_sort_tmp_bytes(n));
  unsigned long long* cnt = (unsigned long long*)B.get(8);
  if (!a || !b || !tmp || !cnt) return SVO_E_OOM;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return SVO_E_HIP;
  int st = SVO_OK;
  const unsigned long long nn = n;
  if (hipMemcpyAsync(a, keys, key_bytes * n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(cnt, &nn, 8, hipMemcpyHostToDevice, s) != hipSuccess) {
    st = SVO_E_HIP;
  } else {
    void* out = nullptr;
    if (key_bytes == 4) {
      uint32_t* r = nullptr;
      launch_radix_sort<uint32_t>((uint32_t*)a, (uint32_t*)b, cnt, n, bits, blocks, tmp, &r, s);
      out = r;
    } else {
      uint64_t* r = nullptr;
      launch_radix_sort<uint64_t>((uint64_t*)a, (uint64_t*)b, cnt, n, bits, blocks, tmp, &r, s);
      out = r;
    }
    if (hipMemcpyAsync(keys, out, key_bytes * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
      st = SVO_E_HIP;
  }
  (void)hipStreamDestroy(s);
  return st;
}

int svo_k_octree_level(int device, const void* sorted, uint64_t n, int key_bytes, int code_bits,
                       void* unique, uint64_t* n_unique, void* parents, uint32_t* first_child,
                       uint64_t* n_parents) {
  if ((key_bytes != 4 && key_bytes != 8) || code_bits < 3 || code_bits >= 8 * key_bytes ||
      !n_unique || !n_parents || n >= (1ull << 32) ||
      (n && (!sorted || !unique || !parents || !first_child)))
    return SVO_E_INVALID;
  *n_unique = *n_parents = 0;
  if (n == 0) return SVO_OK;
  if (hipSetDevice(device) != hipSuccess) return SVO_E_HIP;
  if (key_bytes == 4)
    return octree_level<uint32_t>((const uint32_t*)sorted, n, code_bits, (uint32_t*)unique,
                                  n_unique, (uint32_t*)parents, first_child, n_parents);
  return octree_level<uint64_t>((const uint64_t*)sorted, n, code_bits, (uint64_t*)unique,
                                n_unique, (uint64_t*)parents, first_child, n_parents);
}

int svo_k_ref_vertex_edge(int device, const float* origin, const float* width, uint64_t n_nodes,
                          void* vertices, void* edges) {
  if (n_nodes >= (1ull << 32) || (n_nodes && (!origin || !width || !vertices || !edges)))
    return SVO_E_INVALID;
  if (n_nodes == 0) return SVO_OK;
  if (hipSetDevice(device) != hipSuccess) return SVO_E_HIP;
  DevBufs B;
  float* d_o = (float*)B.get(12 * n_nodes);
  float* d_w = (float*)B.get(4 * n_nodes);
  void* d_v = B.get(sizeof(svo_vertex) * 8 * n_nodes);
  void* d_e = B.get(sizeof(svo_edge) * 12 * n_nodes);
  uint64_t* d_base = (uint64_t*)B.get(16);
  unsigned long long* d_tot = (unsigned long long*)B.get(64);
  if (!d_o || !d_w || !d_v || !d_e || !d_base || !d_tot) return SVO_E_OOM;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return SVO_E_HIP;
  int st = SVO_E_HIP;
  const uint64_t base[2] = {0, n_nodes};  // one mesh: owners are node indices
  if (hipMemcpyAsync(d_o, origin, 12 * n_nodes, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(d_w, width, 4 * n_nodes, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(d_base, base, 16, hipMemcpyHostToDevice, s) == hipSuccess) {
    NodeArrays na{};
    na.origin = d_o;
    na.width = d_w;
    AttrOut out{(svo_vertex*)d_v, (svo_edge*)d_e, nullptr, d_tot};
    launch_ref_attrs(na, n_nodes, d_base, 1, out, s);
    if (hipGetLastError() == hipSuccess &&
        hipMemcpyAsync(vertices, d_v, sizeof(svo_vertex) * 8 * n_nodes, hipMemcpyDeviceToHost,
                       s) == hipSuccess &&
        hipMemcpyAsync(edges, d_e, sizeof(svo_edge) * 12 * n_nodes, hipMemcpyDeviceToHost, s) ==
            hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess)
      st = SVO_OK;
  }
  (void)hipStreamDestroy(s);
  return st;
}

int svo_k_radix_sort_noalloc(int device, void* keys, uint64_t n, int key_bytes, int bits,
                             void* stream) {
  if ((key_bytes != 4 && key_bytes != 8) || bits < 0 || bits > 8 * key_bytes || (n && !keys))
    return SVO_E_INVALID;
  if (n > NOALLOC_SORT_MAX_KEYS) return SVO_E_TOO_LARGE;
  if (n == 0) return SVO_OK;
  // layout: [count | ping-pong keys | digit counts], each on a 256-byte boundary
  const size_t o_keys = 256, o_tmp = o_keys + 8 * NOALLOC_SORT_MAX_KEYS;
  if (o_tmp + radix_sort_tmp_bytes(NOALLOC_SORT_MAX_KEYS) > NOALLOC_SCRATCH_BYTES)
    return SVO_E_TOO_LARGE;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) return SVO_E_HIP;
  int st = SVO_OK;
  void* base = nullptr;
  st = noalloc_scratch(device, &base);
  if (st == SVO_OK) {
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* cnt = (unsigned long long*)base;
    void* pong = (char*)base + o_keys;
    void* tmp = (char*)base + o_tmp;
    const uint32_t blocks = Tuning{}.sort_blocks;
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, s, cnt, (unsigned long long)n);
    void* out = nullptr;
    if (key_bytes == 4) {
      uint32_t* r = nullptr;
      launch_radix_sort<uint32_t>((uint32_t*)keys, (uint32_t*)pong, cnt, n, bits, blocks, tmp,
                                  &r, s);
      out = r;
    } else {
      uint64_t* r = nullptr;
      launch_radix_sort<uint64_t>((uint64_t*)keys, (uint64_t*)pong, cnt, n, bits, blocks, tmp,
                                  &r, s);
      out = r;
    }
    if (hipGetLastError() != hipSuccess) st = SVO_E_HIP;
    // an odd number of passes leaves the keys in the scratch buffer
    if (st == SVO_OK && out != keys &&
        hipMemcpyAsync(keys, out, (size_t)key_bytes * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
      st = SVO_E_HIP;
    if (st == SVO_OK && !stream && hipStreamSynchronize(s) != hipSuccess) st = SVO_E_HIP;
  }
  (void)hipSetDevice(prev);
  return st;
}

int svo_k_voxelize(int device, const float* xyz, uint64_t n_points, const int32_t* tri,
                   uint64_t n_tris, const int32_t* res, uint32_t key_bits, void* codes,
                   uint64_t cap, uint64_t* n_hits) {
  svo_mesh_desc d{xyz, n_points, tri, n_tris};
  svo_params p{};
  for (int a = 0; a < 3; ++a) p.res[a] = res ? res[a] : 0;
  p.key_bits = key_bits;
  int kb = 32;
  int st = validate_params(&d, &p, &kb);
  if (st != SVO_OK) return st;
  if (hipSetDevice(device) != hipSuccess) return SVO_E_HIP;
  int bits = 1;
  for (int a = 0; a < 3; ++a)
    while ((1ll << bits) < (long long)res[a]) ++bits;
  DevBufs B;
  float* dp = (float*)B.get(12 * n_points);
  int32_t* dt = (int32_t*)B.get(12 * n_tris);
  MeshDev* md = (MeshDev*)B.get(sizeof(MeshDev));
  uint64_t* tb = (uint64_t*)B.get(16);
  uint32_t* acc = (uint32_t*)B.get(32);
  MeshParams* par = (MeshParams*)B.get(sizeof(MeshParams));
  unsigned long long* cnt = (unsigned long long*)B.get(64);
  uint32_t* order = (uint32_t*)B.get(order_bytes(n_tris));
  if (!dp || !dt || !md || !tb || !acc || !par || !cnt || !order) return SVO_E_OOM;
  MeshDev h{};
  h.pts = dp;
  h.tri = dt;
  h.n_points = n_points;
  h.n_tris = n_tris;
  for (int a = 0; a < 3; ++a) h.res[a] = res[a];
  const uint64_t htb[2] = {0, n_tris};
  const unsigned long long zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return SVO_E_HIP;
  Tuning tn;
  st = SVO_E_HIP;
  unsigned long long tot[4] = {0, 0, 0, 0};
  if (hipMemcpyAsync(dp, xyz, 12 * n_points, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(dt, tri, 12 * n_tris, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(md, &h, sizeof h, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(tb, htb, 16, hipMemcpyHostToDevice, s) == hipSuccess &&
      hipMemcpyAsync(cnt, zero, 64, hipMemcpyHostToDevice, s) == hipSuccess) {
    launch_bbox(md, 1, n_points, acc, s);
    launch_params(md, 1, acc, par, s);
    launch_classify(md, par, tb, 1, n_tris, tn, cnt, order,
                    (uint8_t*)order + ((4 * n_tris + 255) & ~(uint64_t)255), s);
    if (hipMemcpyAsync(tot, cnt, 32, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess)
      st = SVO_OK;
  }
  MeshParams hp{};
  if (st == SVO_OK) {
    st = SVO_E_HIP;
    if (hipMemcpyAsync(&hp, par, sizeof hp, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess)
      st = hp.status ? SVO_E_INVALID : SVO_OK;
  }
  if (st == SVO_OK) {
    const uint64_t cells = std::max<unsigned long long>(tot[0], 1), n_big = tot[1];
    const size_t kbytes = kb == 64 ? 8 : 4;
    void* keys = B.get(kbytes * cells);
    BigItem* big = (BigItem*)B.get(sizeof(BigItem) * std::max<uint64_t>(n_big, 1));
    if (!keys || !big) {
      st = SVO_E_OOM;
    } else {
      if (kb == 64)
        launch_voxelize<uint64_t>(md, par, tb, 1, n_tris, 3 * bits, tn, (uint64_t*)keys, nullptr,
                                  nullptr, cnt + 2, order, big, cnt + 3, n_big, s);
      else
        launch_voxelize<uint32_t>(md, par, tb, 1, n_tris, 3 * bits, tn, (uint32_t*)keys, nullptr,
                                  nullptr, cnt + 2, order, big, cnt + 3, n_big, s);
      st = SVO_E_HIP;
      if (hipMemcpyAsync(tot, cnt, 32, hipMemcpyDeviceToHost, s) == hipSuccess &&
          hipStreamSynchronize(s) == hipSuccess) {
        const uint64_t h_hits = tot[2];
        if (n_hits) *n_hits = h_hits;
        const uint64_t k = std::min<uint64_t>(h_hits, cap);
        if (k && codes) {
          if (hipMemcpyAsync(codes, keys, kbytes * k, hipMemcpyDeviceToHost, s) == hipSuccess &&
              hipStreamSynchronize(s) == hipSuccess)
            st = SVO_OK;
        } else {
          st = SVO_OK;
        }
      }
    }
  }
  (void)hipStreamDestroy(s);
  return st;
}

}  // extern "C"
