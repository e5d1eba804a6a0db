// svo_device.h -- device-side helpers shared by the gfx950 kernels.
//
// Morton codes: bit i of x, y, z lands on bit 3i, 3i+1, 3i+2.  This is the
// closed form of the reference's byte-LUT encoder (MortonLUT.h:13-233,
// 482-501) and 9-bit-LUT decoder (MortonLUT.h:236-449, 503-529); the golden
// tests check the closed form against those tables.
//
// Keys.  A batch sorts composite keys (mesh << CB) | code, CB = code bits of
// the batch.  A single mesh is the batch of one (mesh bits 0), so the key is
// the plain Morton code.  The parent of a key keeps the mesh bits and shifts
// the code: getParentMorton (SVO.cu:22-24) on the code part.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svo/svo.h"

namespace svo {

using svo_vertex_t = ::svo_vertex;
using svo_edge_t = ::svo_edge;

constexpr uint32_t NONE = 0xFFFFFFFFu;  // UINT_MAX marks absent links (SVO.h:27-33)

// ---------------------------------------------------------------- Morton
__host__ __device__ inline uint32_t spread3_32(uint32_t v) {  // 10 significant bits
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__host__ __device__ inline uint64_t spread3_64(uint64_t v) {  // 21 significant bits
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
__host__ __device__ inline uint64_t compact3_64(uint64_t v) {
  v &= 0x1249249249249249ull;
  v = (v ^ (v >> 2)) & 0x10C30C30C30C30C3ull;
  v = (v ^ (v >> 4)) & 0x100F00F00F00F00Full;
  v = (v ^ (v >> 8)) & 0x1F0000FF0000FFull;
  v = (v ^ (v >> 16)) & 0x1F00000000FFFFull;
  v = (v ^ (v >> 32)) & 0x1FFFFFull;
  return v;
}
__host__ __device__ inline uint32_t compact3_32(uint32_t v) {
  v &= 0x09249249u;
  v = (v ^ (v >> 2)) & 0x030C30C3u;
  v = (v ^ (v >> 4)) & 0x0300F00Fu;
  v = (v ^ (v >> 8)) & 0x030000FFu;
  v = (v ^ (v >> 16)) & 0x000003FFu;
  return v;
}

template <typename K>
__host__ __device__ inline K morton3(uint32_t x, uint32_t y, uint32_t z);
template <>
__host__ __device__ inline uint32_t morton3<uint32_t>(uint32_t x, uint32_t y, uint32_t z) {
  return spread3_32(x) | (spread3_32(y) << 1) | (spread3_32(z) << 2);
}
template <>
__host__ __device__ inline uint64_t morton3<uint64_t>(uint32_t x, uint32_t y, uint32_t z) {
  return spread3_64(x) | (spread3_64(y) << 1) | (spread3_64(z) << 2);
}
template <typename K>
__host__ __device__ inline void demorton3(K m, uint32_t& x, uint32_t& y, uint32_t& z) {
  if (sizeof(K) == 4) {
    x = compact3_32((uint32_t)m);
    y = compact3_32((uint32_t)m >> 1);
    z = compact3_32((uint32_t)m >> 2);
  } else {
    x = (uint32_t)compact3_64((uint64_t)m);
    y = (uint32_t)compact3_64((uint64_t)m >> 1);
    z = (uint32_t)compact3_64((uint64_t)m >> 2);
  }
}

// Streaming 16-byte store for output written once and not read back by the
// same kernel (non-temporal hint: no need to keep the lines in the caches).
typedef uint32_t u32x4_nt __attribute__((ext_vector_type(4)));
__device__ inline void store_stream(uint4* p, uint4 v) {
  u32x4_nt x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, reinterpret_cast<u32x4_nt*>(p));
}

// ---------------------------------------------------------------- keys
template <typename K>
struct KeyLayout {
  int cb;   // code bits
  K cmask;  // (1 << cb) - 1
  __host__ __device__ inline K code(K k) const { return k & cmask; }
  __host__ __device__ inline uint32_t mesh(K k) const { return (uint32_t)(k >> cb); }
  __host__ __device__ inline K parent(K k) const { return (k & ~cmask) | ((k & cmask) >> 3); }
};

// ---------------------------------------------------------------- float ops
// CUDAUtil.h:72-80: non-IEEE min / max (a<b?a:b, a>b?a:b).
__host__ __device__ inline float fmin_r(float a, float b) { return a < b ? a : b; }
__host__ __device__ inline float fmax_r(float a, float b) { return a > b ? a : b; }
// CUDAUtil.h:82-90, 107-110: maxi(lo, mini(f, hi)).
__host__ __device__ inline int clamp_i(int f, int lo, int hi) {
  const int t = f < hi ? f : hi;
  return lo > t ? lo : t;
}

// Order-preserving map of finite floats to uint32 (for atomicMin/Max).
__device__ inline uint32_t float_to_ordered(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ordered_to_float(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(u);
#else
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
#endif
}

// ---------------------------------------------------------------- wave64
__device__ inline uint32_t lane_id() { return __lane_id(); }
// number of set bits of `mask` in lanes below this one
__device__ inline uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// ---------------------------------------------------------------- shared structs
// Per-mesh grid parameters computed on the device from the point bbox.
struct MeshParams {
  float origin[3];  // modelBBox.min after createMeshBBCube (gridOrigin)
  float unit[3];    // unitVoxelSize
  float bmin[3], bmax[3], width[3];
  int32_t res[3];
  uint32_t status;  // 0 or a positive error bit (ERR_*)
};

enum : uint32_t {
  ERR_NONFINITE = 1u << 0,
  ERR_DEGENERATE = 1u << 1,
  ERR_INDEX = 1u << 2,
  ERR_OVERFLOW = 1u << 3,
};

// Per-mesh input descriptor (device copy).
struct MeshDev {
  const float* pts;
  const int32_t* tri;
  uint64_t n_points;
  uint64_t n_tris;
  uint64_t tri_base;  // first global triangle index of this mesh
  uint64_t pt_base;   // first global point index of this mesh
  int32_t res[3];
  uint32_t small_max;  // box cells voxelized by one lane for this mesh (0: the pass's)
};

// Big-triangle work item: a range of (y, z) rows of one triangle's box.
struct BigItem {
  uint32_t tri;       // global triangle index
  uint32_t row0;      // first row (row = y-offset + ny * z-offset)
};

// Per (mesh, level) bookkeeping written by the level kernels.
struct LevelSpan {
  uint32_t elem_first, elem_end;    // range in the level's element array
  uint32_t group_first, group_end;  // range in the level's group array
};

// Per-mesh outcome of the level phase.
struct MeshTreeState {
  uint32_t depth;      // 0 = not finished
  uint32_t root_elem;  // index of the root in its level's element array
  uint32_t n_hits_hi, n_hits_lo;
};

}  // namespace svo
