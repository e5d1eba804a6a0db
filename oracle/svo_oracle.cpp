/*
 * svo_oracle.cpp -- CPU restatement of the Alan-Leo-Wong/GPU-SVO build path.
 *
 * TEST INFRASTRUCTURE ONLY (see svo_oracle.h).  Every function cites the
 * reference file:line it restates.  The structure deliberately follows the
 * reference (dense occupancy marks scanned in index order, per-level
 * head/scan/createNode/createRemainNode steps) rather than the GPU design in
 * gpu-svo_amd/, so that the two are independent implementations.
 *
 * Where the reference has undefined behaviour the oracle applies the rule the
 * product library applies, and says so at the spot:
 *   - a level holding a single code whose low three bits are non-zero, and a
 *     mesh with exactly one voxel (SVO.cu:251 reads an empty scan array;
 *     SVO.cu:257 would write past a one-node array): the single node is placed
 *     at index 0 of its level and its children point to that index;
 *   - empty meshes, out-of-range face indices, non-finite coordinates and a
 *     zero-extent bounding box are rejected with a status code instead of
 *     reading out of bounds or dividing by zero.
 */
#include "svo_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;  // UINT_MAX, SVO.h:27-33

// ---------------------------------------------------------------------------
// Morton LUTs (MortonLUT.h:13-233 encode, 236-449 decode).  The tables are
// generated from their closed form: byte bit i of axis a lands on bit 3i+a;
// the decode tables gather bits a, a+3, a+6 of a 9-bit chunk.  The golden
// tests compare these generated tables with the values printed in the
// reference header.
// ---------------------------------------------------------------------------
struct Luts {
  uint32_t enc[3][256];
  uint8_t dec[3][512];
  int16_t nb_parent[8][27];
  int16_t nb_child[8][27];
  int16_t vert_shared[8][8];
  int16_t edge_shared[12][4];
  Luts() {
    for (int a = 0; a < 3; ++a) {
      for (uint32_t v = 0; v < 256; ++v) {
        uint32_t r = 0;
        for (int i = 0; i < 8; ++i) r |= ((v >> i) & 1u) << (3 * i + a);
        enc[a][v] = r;
      }
      for (uint32_t v = 0; v < 512; ++v) {
        uint32_t r = 0;
        for (int i = 0; i < 3; ++i) r |= ((v >> (3 * i + a)) & 1u) << i;
        dec[a][v] = (uint8_t)r;
      }
    }
    // Neighbor LUTs (MortonLUT.h:531-597): slot s = (dx+1) + 3(dy+1) + 9(dz+1);
    // the neighbour of child k in direction d is child ((k+d) mod 2 per axis)
    // of the parent-level neighbour in direction floor((k+d)/2).
    for (int k = 0; k < 8; ++k) {
      const int kc[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};
      for (int s = 0; s < 27; ++s) {
        const int d[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
        int ps = 0, ch = 0;
        const int mul[3] = {1, 3, 9};
        for (int a = 0; a < 3; ++a) {
          const int c = kc[a] + d[a];            // -1 .. 2
          const int p = (c < 0) ? -1 : (c >> 1);  // -1, 0, 0, 1
          ps += (p + 1) * mul[a];
          ch |= (c & 1) << a;
        }
        nb_parent[k][s] = (int16_t)ps;
        nb_child[k][s] = (int16_t)ch;
      }
    }
    // Cells sharing corner i (SVO.cu:567-584): corner offset b in {0,1}^3;
    // cells at d in {b-1, b} per axis, listed with x fastest.
    for (int i = 0; i < 8; ++i) {
      const int b[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
      int j = 0;
      for (int dz = b[2] - 1; dz <= b[2]; ++dz)
        for (int dy = b[1] - 1; dy <= b[1]; ++dy)
          for (int dx = b[0] - 1; dx <= b[0]; ++dx)
            vert_shared[i][j++] = (int16_t)((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1));
    }
    // Cells sharing edge i (SVO.cu:610-627), edges in the order of
    // SVO.cu:638-663.  Each edge: fixed axis offsets for the two axes it does
    // not run along; cells span {o-1, o} on those two axes, ascending slot.
    // Edge = (axis it runs along, offset on the other two axes x,y,z with -1
    // marking the running axis).
    const int edge_def[12][3] = {
        {0, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, 0, 0},   // z = 0 face ring
        {0, -1, 1}, {-1, 1, 1}, {1, -1, 1}, {-1, 0, 1},   // z = 1 face ring
        {0, 0, -1}, {0, 1, -1}, {1, 1, -1}, {1, 0, -1}};  // vertical edges
    for (int e = 0; e < 12; ++e) {
      int slots[4], n = 0;
      for (int s = 0; s < 27; ++s) {
        const int d[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
          if (edge_def[e][a] < 0) {
            ok = ok && d[a] == 0;
          } else {
            ok = ok && (d[a] == edge_def[e][a] - 1 || d[a] == edge_def[e][a]);
          }
        }
        if (ok) slots[n++] = s;
      }
      for (int j = 0; j < 4; ++j) edge_shared[e][j] = (int16_t)slots[j];
    }
  }
};
const Luts& luts() {
  static Luts L;
  return L;
}

// mortonEncode_LUT(uint16_t, uint16_t, uint16_t), MortonLUT.h:482-501:
// two byte-wide lookups combined with a 24-bit shift in uint32 arithmetic.
inline uint32_t encode_u16(uint32_t x, uint32_t y, uint32_t z) {
  const Luts& L = luts();
  uint32_t answer = 0;
  for (unsigned i = 2; i > 0; --i) {
    const unsigned shift = (i - 1) * 8;
    answer = answer << 24 |
             (L.enc[2][(z >> shift) & 0xFF] | L.enc[1][(y >> shift) & 0xFF] |
              L.enc[0][(x >> shift) & 0xFF]);
  }
  return answer;
}

// 64-bit widening of the same byte-wise interleave (three bytes per axis,
// 21 significant bits), used for resolutions the uint32 codes cannot hold.
inline uint64_t encode_u64(uint32_t x, uint32_t y, uint32_t z) {
  const Luts& L = luts();
  uint64_t answer = 0;
  for (unsigned i = 3; i > 0; --i) {
    const unsigned shift = (i - 1) * 8;
    answer = answer << 24 |
             (uint64_t)(L.enc[2][(z >> shift) & 0xFF] | L.enc[1][(y >> shift) & 0xFF] |
                        L.enc[0][(x >> shift) & 0xFF]);
  }
  return answer;
}

// morton3D_32_decode / morton3D_DecodeCoord_LUT256, MortonLUT.h:503-529:
// nine-bit chunks, seven of them (the width of uint_fast32_t on LP64).
inline void decode(uint64_t m, uint32_t& x, uint32_t& y, uint32_t& z) {
  const Luts& L = luts();
  uint64_t a[3] = {0, 0, 0};
  for (unsigned i = 0; i < 7; ++i)
    for (int k = 0; k < 3; ++k) a[k] |= (uint64_t)L.dec[k][(m >> (i * 9)) & 0x1FF] << (3 * i);
  x = (uint32_t)a[0];
  y = (uint32_t)a[1];
  z = (uint32_t)a[2];
}

// getParentMorton, SVO.cu:22-24 ((m >> 3) & 0x0FFFFFFF for uint32 codes).
inline uint64_t parent_code(uint64_t m, int key_bits) {
  return key_bits == 32 ? ((m >> 3) & 0x0FFFFFFFull) : (m >> 3);
}

// CUDAUtil.h:72-80: non-IEEE min / max.
inline float fmin_r(float a, float b) { return a < b ? a : b; }
inline float fmax_r(float a, float b) { return a > b ? a : b; }
// CUDAUtil.h:82-90, 107-110.
inline int clamp_i(int f, int a, int b) {
  const int lo = f < b ? f : b;
  return a > lo ? a : lo;
}

struct V3 {
  float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// Arithmetic variant, for measuring how sensitive the voxel set is to what
// cannot be pinned here (the default, 0, is the pinned restatement):
//   bit 0  length-3 sums as (a0 + a1) + a2 instead of Eigen's a0 + (a1 + a2)
//   bit 1  products feeding an add contracted to fused multiply-adds, as
//          nvcc does by default (-fmad=true) in dots and the cross product
int g_arith = 0;

// Eigen 3.4 fixed-size redux splits a length-3 sum as a0 + (a1 + a2).
inline float dot3(V3 a, V3 b) {
  if (__builtin_expect(g_arith == 0, 1)) return a.x * b.x + (a.y * b.y + a.z * b.z);
  if (g_arith & 2) {
    if (g_arith & 1) return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x));
    return std::fma(a.x, b.x, std::fma(a.y, b.y, a.z * b.z));
  }
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
inline float dot2(float a0, float a1, float b0, float b1) {
  if (__builtin_expect(g_arith & 2, 0)) return std::fma(a0, b0, a1 * b1);
  return a0 * b0 + a1 * b1;
}
inline float cross_c(float a, float b, float c, float d) {
  if (__builtin_expect(g_arith & 2, 0)) return std::fma(a, b, -(c * d));
  return a * b - c * d;
}

struct Grid {
  float origin[3];  // modelBBox.min
  float unit[3];    // unitVoxelSize
  float bmin[3], bmax[3], width[3];
  int res[3];
};

// BaseModel.h:64-73 (min/max over all points) and createMeshBBCube,
// BaseModel.h:24-49; then SVO.cu:348-350 for the voxel size.
int make_grid(const float* P, uint64_t np, const int32_t* res, Grid& g) {
  float mn[3], mx[3];
  for (int a = 0; a < 3; ++a) mn[a] = mx[a] = P[a];
  for (uint64_t i = 0; i < np; ++i)
    for (int a = 0; a < 3; ++a) {
      const float v = P[3 * i + a];
      if (!std::isfinite(v)) return ORC_E_INVALID;
      if (v < mn[a]) mn[a] = v;
      if (v > mx[a]) mx[a] = v;
    }
  float lengths[3], amin[3], amax[3];
  for (int a = 0; a < 3; ++a) {
    lengths[a] = mx[a] - mn[a];
    amin[a] = mn[a];
    amax[a] = mx[a];
  }
  const float max_length = fmax_r(lengths[0], fmax_r(lengths[1], lengths[2]));
  if (!(max_length > 0.0f)) return ORC_E_INVALID;  // reference divides by zero later
  for (int a = 0; a < 3; ++a) {
    if (max_length == lengths[a]) continue;
    const float delta = max_length - lengths[a];
    amin[a] = mn[a] - (delta / 2.0f);
    amax[a] = mx[a] + (delta / 2.0f);
  }
  for (int a = 0; a < 3; ++a) {
    const float eps = (amax[a] - amin[a]) / 10001.0f;
    amin[a] -= eps;
    amax[a] += eps;
  }
  for (int a = 0; a < 3; ++a) {
    g.bmin[a] = amin[a];
    g.bmax[a] = amax[a];
    g.width[a] = amax[a] - amin[a];
    g.origin[a] = amin[a];
    g.res[a] = res[a];
    g.unit[a] = g.width[a] / (float)res[a];
  }
  return ORC_OK;
}

// surfaceVoxelize, SVO.cu:37-188, for one triangle.  emit(x, y, z) is called
// for every grid cell that passes the plane and the three projection tests.
template <class Emit>
inline void voxelize_triangle(const Grid& g, V3 t0, V3 t1, V3 t2, Emit&& emit) {
  const V3 origin = {g.origin[0], g.origin[1], g.origin[2]};
  const V3 u = {g.unit[0], g.unit[1], g.unit[2]};
  const V3 v0 = sub(t0, origin), v1 = sub(t1, origin), v2 = sub(t2, origin);  // 58-63
  const V3 e0 = sub(v1, v0), e1 = sub(v2, v1), e2 = sub(v0, v2);              // 65-67
  // e0.cross(e1).normalized(), 69
  V3 n = {cross_c(e0.y, e1.z, e0.z, e1.y), cross_c(e0.z, e1.x, e0.x, e1.z),
          cross_c(e0.x, e1.y, e0.y, e1.x)};
  const float zz = dot3(n, n);
  if (zz > 0.0f) {
    const float s = std::sqrt(zz);
    n = {n.x / s, n.y / s, n.z / s};
  }
  // 73-85: bounding box in grid cells.
  const V3 bmin = {fmin_r(v0.x, fmin_r(v1.x, v2.x)), fmin_r(v0.y, fmin_r(v1.y, v2.y)),
                   fmin_r(v0.z, fmin_r(v1.z, v2.z))};
  const V3 bmax = {fmax_r(v0.x, fmax_r(v1.x, v2.x)), fmax_r(v0.y, fmax_r(v1.y, v2.y)),
                   fmax_r(v0.z, fmax_r(v1.z, v2.z))};
  const int gx0 = clamp_i((int)(bmin.x / u.x), 0, g.res[0] - 1);
  const int gy0 = clamp_i((int)(bmin.y / u.y), 0, g.res[1] - 1);
  const int gz0 = clamp_i((int)(bmin.z / u.z), 0, g.res[2] - 1);
  const int gx1 = clamp_i((int)(bmax.x / u.x), 0, g.res[0] - 1);
  const int gy1 = clamp_i((int)(bmax.y / u.y), 0, g.res[1] - 1);
  const int gz1 = clamp_i((int)(bmax.z / u.z), 0, g.res[2] - 1);
  // 88-93: plane test constants.
  const V3 c = {n.x > 0.0f ? u.x : 0.0f, n.y > 0.0f ? u.y : 0.0f, n.z > 0.0f ? u.z : 0.0f};
  const float d1 = dot3(n, sub(c, v0));
  const float d2 = dot3(n, sub(sub(u, c), v0));
  // 96-149: projection test constants.  ne[plane][edge] = 2-vector.
  float ne[3][3][2], de[3][3];
  const V3 ev[3] = {e0, e1, e2};
  const V3 vv[3] = {v0, v1, v2};
  for (int i = 0; i < 3; ++i) {
    // XY: (-e.y, e.x), flipped when n.z < 0
    ne[0][i][0] = -1.0f * ev[i].y;
    ne[0][i][1] = ev[i].x;
    // YZ: (-e.z, e.y), flipped when n.x < 0
    ne[1][i][0] = -1.0f * ev[i].z;
    ne[1][i][1] = ev[i].y;
    // ZX: (-e.x, e.z), flipped when n.y < 0
    ne[2][i][0] = -1.0f * ev[i].x;
    ne[2][i][1] = ev[i].z;
  }
  const bool flip[3] = {n.z < 0.0f, n.x < 0.0f, n.y < 0.0f};
  for (int p = 0; p < 3; ++p)
    if (flip[p])
      for (int i = 0; i < 3; ++i) {
        ne[p][i][0] = -ne[p][i][0];
        ne[p][i][1] = -ne[p][i][1];
      }
  for (int i = 0; i < 3; ++i) {
    // plane axes (a, b): XY = (x, y) with units (u.x, u.y); YZ = (y, z); ZX = (z, x)
    const float va[3] = {vv[i].x, vv[i].y, vv[i].z};
    const float vb[3] = {vv[i].y, vv[i].z, vv[i].x};
    const float ua[3] = {u.x, u.y, u.z};
    const float ub[3] = {u.y, u.z, u.x};
    for (int p = 0; p < 3; ++p) {
      de[p][i] = ((-1.0f * dot2(ne[p][i][0], ne[p][i][1], va[p], vb[p])) +
                  fmax_r(0.0f, ua[p] * ne[p][i][0])) +
                 fmax_r(0.0f, ub[p] * ne[p][i][1]);
    }
  }
  // 152-185: test every cell of the clamped box.
  for (int z = gz0; z <= gz1; ++z)
    for (int y = gy0; y <= gy1; ++y)
      for (int x = gx0; x <= gx1; ++x) {
        const V3 p = {(float)x * u.x, (float)y * u.y, (float)z * u.z};
        const float t = dot3(n, p);
        if ((t + d1) * (t + d2) > 0.0f) continue;
        const float pa[3] = {p.x, p.y, p.z};
        const float pb[3] = {p.y, p.z, p.x};
        bool ok = true;
        for (int q = 0; q < 3 && ok; ++q)
          for (int i = 0; i < 3; ++i)
            if ((dot2(ne[q][i][0], ne[q][i][1], pa[q], pb[q]) + de[q][i]) < 0.0f) {
              ok = false;
              break;
            }
        if (!ok) continue;
        emit((uint32_t)x, (uint32_t)y, (uint32_t)z);
      }
}

int validate(const float* P, uint64_t np, const int32_t* T, uint64_t nt, const int32_t* res,
             int key_bits) {
  if (np == 0 || nt == 0) return ORC_E_EMPTY;
  if (key_bits != 32 && key_bits != 64) return ORC_E_INVALID;
  for (int a = 0; a < 3; ++a) {
    if (res[a] < 1) return ORC_E_INVALID;
    if (key_bits == 32 && res[a] > 1024) return ORC_E_RES_LIMIT;
    if (key_bits == 64 && res[a] > (1 << 21)) return ORC_E_RES_LIMIT;
  }
  for (uint64_t i = 0; i < 3 * nt; ++i)
    if (T[i] < 0 || (uint64_t)T[i] >= np) return ORC_E_INVALID;
  (void)P;
  return ORC_OK;
}

// Level-1 voxel set: the reference scatters code|flag into a dense array
// indexed by code (SVO.cu:180-182) and compacts it in index order
// (SVO.cu:386-405), which yields the sorted unique codes.  The oracle keeps the
// dense occupancy as a bitmap for uint32 codes, and sorts for uint64 codes.
int voxelize_all(const float* P, const int32_t* T, uint64_t nt, const Grid& g, int key_bits,
                 int n_threads, std::vector<uint64_t>& out) {
  out.clear();
  auto tri = [&](uint64_t t, int k) {
    const int32_t i = T[3 * t + k];
    return V3{P[3 * i], P[3 * i + 1], P[3 * i + 2]};
  };
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
  n_threads = 1;
#endif
  if (key_bits == 32) {
    // gridCNodeSize = morton(res-1) + 1, SVO.cu:343-345
    const uint64_t G =
        (uint64_t)encode_u16(g.res[0] - 1, g.res[1] - 1, g.res[2] - 1) + 1;
    std::vector<uint64_t> bits((G + 63) / 64, 0);
    uint64_t* B = bits.data();
#pragma omp parallel for schedule(dynamic, 256) num_threads(n_threads)
    for (int64_t t = 0; t < (int64_t)nt; ++t) {
      voxelize_triangle(g, tri(t, 0), tri(t, 1), tri(t, 2), [&](uint32_t x, uint32_t y, uint32_t z) {
        const uint32_t code = encode_u16(x, y, z);
        __atomic_fetch_or(&B[code >> 6], 1ull << (code & 63), __ATOMIC_RELAXED);
      });
    }
    for (uint64_t w = 0; w < bits.size(); ++w) {
      uint64_t m = bits[w];
      while (m) {
        const int b = __builtin_ctzll(m);
        out.push_back(w * 64 + b);
        m &= m - 1;
      }
    }
  } else {
    std::vector<std::vector<uint64_t>> parts(n_threads);
#pragma omp parallel num_threads(n_threads)
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      std::vector<uint64_t>& mine = parts[tid];
#pragma omp for schedule(dynamic, 256)
      for (int64_t t = 0; t < (int64_t)nt; ++t)
        voxelize_triangle(g, tri(t, 0), tri(t, 1), tri(t, 2),
                          [&](uint32_t x, uint32_t y, uint32_t z) { mine.push_back(encode_u64(x, y, z)); });
      std::sort(mine.begin(), mine.end());
      mine.erase(std::unique(mine.begin(), mine.end()), mine.end());
    }
    for (auto& p : parts) out.insert(out.end(), p.begin(), p.end());
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }
  return out.empty() ? ORC_E_EMPTY : ORC_OK;
}

struct VertRec {
  float p[3];
  uint32_t owner;
};
struct EdgeRec {
  float a[3];
  float b[3];
  uint32_t owner;
};
static_assert(sizeof(VertRec) == 16, "vertex record is 16 bytes");
static_assert(sizeof(EdgeRec) == 28, "edge record is 28 bytes");

struct Tree {
  oracle_info info;
  std::vector<uint64_t> voxels;
  std::vector<uint64_t> code;
  std::vector<uint8_t> is_leaf;
  std::vector<float> origin;
  std::vector<float> width;
  std::vector<uint32_t> parent, childs, neighbors;
  std::vector<VertRec> verts;
  std::vector<EdgeRec> edges, faces;
  std::vector<uint64_t> level_nodes;
};

// One level of the tree as the reference keeps it before appending it to the
// global array (SVO.cu:440-471).
struct LevelNodes {
  std::vector<uint64_t> code;
  std::vector<uint8_t> is_leaf, filled;
  std::vector<float> origin, width;
  std::vector<uint32_t> childs;
};

// createNode_1 / createNode_2 / createRemainNode field writes.
inline void set_node_geometry(LevelNodes& L, uint64_t a, uint64_t code, const Grid& g, float w) {
  uint32_t x, y, z;
  decode(code, x, y, z);
  L.code[a] = code;
  // gridOrigin + width * Vector3f(x, y, z): per axis o + (w * c), SVO.cu:262
  L.origin[3 * a + 0] = g.origin[0] + w * (float)x;
  L.origin[3 * a + 1] = g.origin[1] + w * (float)y;
  L.origin[3 * a + 2] = g.origin[2] + w * (float)z;
  L.width[a] = w;
  L.filled[a] = 1;
}

// SparseVoxelOctree::create level loop, SVO.cu:383-495, restated over sorted
// unique code vectors instead of dense flag arrays.
void build_levels(const Grid& g, int key_bits, const std::vector<uint64_t>& s1, Tree& T) {
  std::vector<uint64_t> pact = s1;  // compacted codes of the current level
  std::vector<uint64_t> esum;       // d_esumTreeNodesArray
  std::vector<uint64_t> sum;        // d_sumTreeNodesArray (persists across levels)
  float w = g.unit[0];              // unitNodeWidth = unitVoxelSize.x(), SVO.cu:351
  int depth = 0;
  while (true) {
    const uint64_t numC = pact.size();  // 391
    if (!numC) break;
    ++depth;  // 397
    uint64_t numNodes = 1;  // 422
    std::vector<uint64_t> next;
    if (numC > 1) {  // 423-438
      std::vector<uint64_t> nNodes(numC, 0);
      for (uint64_t i = 1; i < numC; ++i)  // cpNumNodes, 221-235
        nNodes[i] = parent_code(pact[i], key_bits) == parent_code(pact[i - 1], key_bits) ? 0 : 8;
      sum.assign(numC, 0);
      uint64_t run = 0;
      for (uint64_t i = 0; i < numC; ++i) sum[i] = (run += nNodes[i]);  // inclusive_scan, 434
      numNodes = sum[numC - 1] + 8;                                     // 436
      // parents flagged in the dense parent array (429, 231), compacted in
      // index order at the next iteration
      for (uint64_t i = 0; i < numC; ++i) {
        const uint64_t p = parent_code(pact[i], key_bits);
        if (next.empty() || next.back() != p) next.push_back(p);
      }
    }
    T.level_nodes.push_back(numNodes);
    LevelNodes L;  // SVONode() defaults, 442
    L.code.assign(numNodes, 0);
    L.is_leaf.assign(numNodes, 1);
    L.filled.assign(numNodes, 0);
    L.origin.assign(3 * numNodes, 0.0f);
    L.width.assign(numNodes, 0.0f);
    L.childs.assign(8 * numNodes, NONE);
    const uint64_t level_off = esum.empty() ? 0 : esum.back();
    std::vector<uint64_t> beg(numC);
    if (depth < 2) {
      esum.push_back(0);  // 452
    }
    const uint64_t preDepth = esum.back();                          // this level's offset
    const uint64_t preChild = esum.size() >= 2 ? esum[esum.size() - 2] : 0;
    for (uint64_t i = 0; i < numC; ++i) {  // createNode_1 237-267 / createNode_2 269-309
      const uint64_t code = pact[i];
      const uint64_t key = code & 7;
      // Single-code level: the reference reads a stale or empty scan array
      // and writes at index `key` of a one-node array; place it at index 0.
      const uint64_t address = (numC == 1) ? 0 : sum[i] + key;
      set_node_geometry(L, address, code, g, w);
      beg[i] = (code / 8) * 8;
      if (depth >= 2) {
        L.is_leaf[address] = 0;
        for (int c = 0; c < 8; ++c) {
          L.childs[8 * address + c] = (uint32_t)(preChild + i * 8 + c);
          T.parent[preChild + i * 8 + c] = (uint32_t)(preDepth + address);
        }
      }
    }
    (void)level_off;
    beg.erase(std::unique(beg.begin(), beg.end()), beg.end());  // 461-462
    for (uint64_t t = 0; t < numNodes; ++t) {                  // createRemainNode 311-339
      if (L.code[t] == 0) {
        const uint64_t code = beg[t / 8] + (t & 7);
        const uint8_t keep_leaf = L.is_leaf[t];
        set_node_geometry(L, t, code, g, w);
        L.is_leaf[t] = keep_leaf;
      }
    }
    // append the level (d_SVONodeArray.insert, 471) and extend per-node arrays
    T.code.insert(T.code.end(), L.code.begin(), L.code.end());
    T.is_leaf.insert(T.is_leaf.end(), L.is_leaf.begin(), L.is_leaf.end());
    T.origin.insert(T.origin.end(), L.origin.begin(), L.origin.end());
    T.width.insert(T.width.end(), L.width.begin(), L.width.end());
    T.childs.insert(T.childs.end(), L.childs.begin(), L.childs.end());
    T.parent.resize(T.code.size(), NONE);
    esum.push_back(numNodes + esum.back());  // 473
    // 476-494: the next dense size comes from the largest flagged parent; the
    // loop ends once a level holds a single node.
    if (numNodes / 8 == 0) break;
    w *= 2.0;  // 483 (float *= double: exact doubling)
    pact.swap(next);
  }
  T.info.depth = depth;
}

// constructNodeNeighbors, SVO.cu:548-565.  Reference semantics: findNeighbors
// <false> writes a local copy (535-542), so only the root's slot 13 is stored
// (528-529).  Intended semantics: the top-down LUT rule of 536-542 applied to
// the stored array.
void build_neighbors(int semantics, Tree& T) {
  const uint64_t N = T.code.size();
  T.neighbors.assign(27 * N, NONE);
  const int depth = T.info.depth;
  if (depth < 2) return;  // 554
  const uint64_t root = N - 1;
  T.neighbors[27 * root + 13] = (uint32_t)root;
  if (semantics != ORC_SEM_INTENDED) return;
  const Luts& L = luts();
  std::vector<uint64_t> off(depth + 1, 0);
  for (int l = 0; l < depth; ++l) off[l + 1] = off[l] + T.level_nodes[l];
  for (int l = depth - 2; l >= 0; --l) {  // 558-563, top-down
    for (uint64_t t = off[l]; t < off[l + 1]; ++t) {
      const uint32_t p = T.parent[t];
      const int key = (int)(T.code[t] & 7);
      for (int s = 0; s < 27; ++s) {
        const uint32_t pn = T.neighbors[27 * (uint64_t)p + L.nb_parent[key][s]];
        T.neighbors[27 * t + s] =
            (pn != NONE) ? T.childs[8 * (uint64_t)pn + L.nb_child[key][s]] : NONE;
      }
    }
  }
}

inline bool eq3(const float* a, const float* b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }

// determineNodeVertex / determineNodeEdge / uniqueVert / uniqueEdge,
// SVO.cu:586-728, and the intended owner-based de-duplication.
void build_vertex_edge(int semantics, Tree& T) {
  const Luts& L = luts();
  const uint64_t N = T.code.size();
  const uint32_t* nb = T.neighbors.data();
  // edge endpoint offsets in units of width, SVO.cu:638-663
  static const int E[12][2][3] = {
      {{0, 0, 0}, {0, 1, 0}}, {{0, 1, 0}, {1, 1, 0}}, {{1, 1, 0}, {1, 0, 0}}, {{1, 0, 0}, {0, 0, 0}},
      {{0, 0, 1}, {0, 1, 1}}, {{0, 1, 1}, {1, 1, 1}}, {{1, 1, 1}, {1, 0, 1}}, {{1, 0, 1}, {0, 0, 1}},
      {{0, 0, 0}, {0, 0, 1}}, {{0, 1, 0}, {0, 1, 1}}, {{1, 1, 0}, {1, 1, 1}}, {{1, 0, 0}, {1, 0, 1}}};
  // origin + Vector3f(...) adds to every component; the bare `origin` term
  // (offset 0,0,0) is used without an addition (SVO.cu:640, 645, 656).
  auto endpoint = [&](uint64_t t, const int* o, float* out) {
    const float w = T.width[t];
    const bool bare = !o[0] && !o[1] && !o[2];
    for (int a = 0; a < 3; ++a)
      out[a] = bare ? T.origin[3 * t + a] : T.origin[3 * t + a] + (o[a] ? w : 0.0f);
  };
  T.verts.clear();
  T.edges.clear();
  T.faces.clear();
  if (semantics != ORC_SEM_INTENDED) {
    std::vector<VertRec> v(8 * N);
    std::vector<EdgeRec> e(12 * N);
    for (uint64_t t = 0; t < N; ++t) {
      const float w = T.width[t];
      for (int i = 0; i < 8; ++i) {
        uint32_t x, y, z;
        decode((uint64_t)i, x, y, z);  // 597
        VertRec& r = v[8 * t + i];
        r.p[0] = w * (float)x;  // corner = width * (x, y, z); origin unused (598)
        r.p[1] = w * (float)y;
        r.p[2] = w * (float)z;
        uint32_t idx = (uint32_t)t;  // 600-603: `morton` never updated -> last valid wins
        for (int j = 0; j < 8; ++j)
          if (nb[27 * t + L.vert_shared[i][j]] < NONE) idx = nb[27 * t + L.vert_shared[i][j]];
        r.owner = idx;
      }
      for (int i = 0; i < 12; ++i) {
        EdgeRec& r = e[12 * t + i];
        endpoint(t, E[i][0], r.a);
        endpoint(t, E[i][1], r.b);
        uint32_t idx = (uint32_t)t;  // 669-672
        for (int j = 0; j < 4; ++j)
          if (nb[27 * t + L.edge_shared[i][j]] < NONE) idx = nb[27 * t + L.edge_shared[i][j]];
        r.owner = idx;
      }
    }
    // thrust::unique: drop an element equal to its predecessor (710-711, 719-720)
    for (uint64_t i = 0; i < v.size(); ++i)
      if (i == 0 || !eq3(v[i].p, v[i - 1].p)) T.verts.push_back(v[i]);
    for (uint64_t i = 0; i < e.size(); ++i) {
      bool dup = false;
      if (i > 0) {
        const EdgeRec &a = e[i - 1], &b = e[i];
        dup = (eq3(a.a, b.a) && eq3(a.b, b.b)) || (eq3(a.a, b.b) && eq3(a.b, b.a));
      }
      if (!dup) T.edges.push_back(e[i]);
    }
    return;
  }
  // Intended: every vertex / edge / face shared by same-level cells is owned
  // by the sharing node with the smallest index (= smallest Morton code in
  // the level) and emitted once, by its owner, in node-major order.
  static const int FACE_SLOT[6] = {12, 14, 10, 16, 4, 22};
  static const int F[6][2][3] = {{{0, 0, 0}, {0, 1, 1}}, {{1, 0, 0}, {1, 1, 1}},
                                 {{0, 0, 0}, {1, 0, 1}}, {{0, 1, 0}, {1, 1, 1}},
                                 {{0, 0, 0}, {1, 1, 0}}, {{0, 0, 1}, {1, 1, 1}}};
  for (uint64_t t = 0; t < N; ++t) {
    for (int i = 0; i < 8; ++i) {
      uint32_t owner = (uint32_t)t;
      for (int j = 0; j < 8; ++j) owner = std::min(owner, nb[27 * t + L.vert_shared[i][j]]);
      if (owner != t) continue;
      const int o[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
      VertRec r;
      endpoint(t, o, r.p);
      r.owner = owner;
      T.verts.push_back(r);
    }
    for (int i = 0; i < 12; ++i) {
      uint32_t owner = (uint32_t)t;
      for (int j = 0; j < 4; ++j) owner = std::min(owner, nb[27 * t + L.edge_shared[i][j]]);
      if (owner != t) continue;
      EdgeRec r;
      endpoint(t, E[i][0], r.a);
      endpoint(t, E[i][1], r.b);
      r.owner = owner;
      T.edges.push_back(r);
    }
    for (int f = 0; f < 6; ++f) {
      const uint32_t owner = std::min((uint32_t)t, nb[27 * t + FACE_SLOT[f]]);
      if (owner != t) continue;
      EdgeRec r;
      endpoint(t, F[f][0], r.a);
      endpoint(t, F[f][1], r.b);
      r.owner = owner;
      T.faces.push_back(r);
    }
  }
}

int build_tree(const float* P, uint64_t np, const int32_t* Ti, uint64_t nt, const int32_t* res,
               int semantics, int key_bits, int n_threads, Tree& T) {
  std::memset(&T.info, 0, sizeof(T.info));
  T.info.key_bits = key_bits;
  T.info.semantics = semantics;
  for (int a = 0; a < 3; ++a) T.info.res[a] = res[a];
  int st = validate(P, np, Ti, nt, res, key_bits);
  if (st) return st;
  Grid g;
  st = make_grid(P, np, res, g);
  if (st) return st;
  for (int a = 0; a < 3; ++a) {
    T.info.bbox_min[a] = g.bmin[a];
    T.info.bbox_max[a] = g.bmax[a];
    T.info.bbox_width[a] = g.width[a];
    T.info.unit[a] = g.unit[a];
  }
  const auto t0 = std::chrono::steady_clock::now();
  st = voxelize_all(P, Ti, nt, g, key_bits, n_threads, T.voxels);
  if (st) return st;
  build_levels(g, key_bits, T.voxels, T);
  build_neighbors(semantics, T);
  build_vertex_edge(semantics, T);
  const auto t1 = std::chrono::steady_clock::now();
  T.info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
  T.info.n_voxels = T.voxels.size();
  T.info.n_nodes = T.code.size();
  T.info.n_vertices = T.verts.size();
  T.info.n_edges = T.edges.size();
  T.info.n_faces = T.faces.size();
  for (size_t l = 0; l < T.level_nodes.size() && l < 32; ++l) T.info.level_nodes[l] = T.level_nodes[l];
  return ORC_OK;
}

// write_vertex / write_line / write_cube, IO.h:8-90, and writeTree,
// SVO.cu:737-753: ostream output of float is printf "%g" of the value
// promoted to double; every line ends with std::endl.
// write_cube, IO.h:24-93, for each of N cubes in order.
int write_cubes(const float* origin, const float* width, uint64_t N, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return ORC_E_IO;
  std::vector<char> buf(1 << 20);
  std::setvbuf(f, buf.data(), _IOFBF, buf.size());
  uint64_t base = 0;
  for (uint64_t t = 0; t < N; ++t) {
    const float* o = &origin[3 * t];
    const float w = width[t];
    // v1..v8 of IO.h:133-140: origin + (0|w per axis)
    static const int C[8][3] = {{0, 1, 1}, {0, 0, 1}, {0, 1, 0}, {0, 0, 0},
                                {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    for (int k = 0; k < 8; ++k) {
      float v[3];
      const bool bare = k == 3;  // v4 = nodeOrigin, IO.h:136
      for (int a = 0; a < 3; ++a) v[a] = bare ? o[a] : o[a] + (C[k][a] ? w : 0.0f);
      std::fprintf(f, "v %g %g %g\n", (double)v[0], (double)v[1], (double)v[2]);
    }
    static const int Lp[12][2] = {{1, 2}, {2, 7}, {7, 8}, {8, 1}, {3, 4}, {4, 5},
                                  {5, 6}, {6, 3}, {3, 1}, {4, 2}, {5, 7}, {6, 8}};
    for (int k = 0; k < 12; ++k)
      std::fprintf(f, "l %d %d\n", (int)(base + Lp[k][0]), (int)(base + Lp[k][1]));
    base += 8;
  }
  return std::fclose(f) == 0 ? ORC_OK : ORC_E_IO;
}

// writeTree, SVO.cu:737-753: every node in array order.
int write_tree(const Tree& T, const char* path) {
  return write_cubes(T.origin.data(), T.width.data(), T.code.size(), path);
}

// writeVoxel, SVO.cu:755-777: one cube per voxel of S1 in sorted order at
// modelBBox.min + unitNodeWidth * (x, y, z), edge unitNodeWidth (= unit.x).
int write_voxels(const Tree& T, const char* path) {
  const uint64_t U = T.voxels.size();
  const float w = T.info.unit[0];
  std::vector<float> org(3 * U), wid(U, w);
  for (uint64_t i = 0; i < U; ++i) {
    uint32_t c[3];
    decode(T.voxels[i], c[0], c[1], c[2]);
    for (int a = 0; a < 3; ++a) org[3 * i + a] = T.info.bbox_min[a] + w * (float)c[a];
  }
  return write_cubes(org.data(), wid.data(), U, path);
}

// 168-byte SVONode record, SVO.h:20-35.
struct SVONodeRec {
  uint32_t mortonCode;
  uint8_t isLeaf;
  uint8_t pad[3];
  float origin[3];
  float width;
  uint32_t parent;
  uint32_t childs[8];
  uint32_t neighbors[27];
};
static_assert(sizeof(SVONodeRec) == 168, "SVONode is 168 bytes");

}  // namespace

extern "C" {

int oracle_load_obj(const char* path, float** points, uint64_t* n_points, int32_t** tris,
                    uint64_t* n_tris) {
  // BaseModel::loadOBJ, BaseModel.cpp:3-25 (sscanf in place of MSVC sscanf_s).
  std::ifstream in(path);
  if (!in) return ORC_E_IO;  // 4-8
  std::vector<float> P;
  std::vector<int32_t> F;
  float x, y, z;
  int f0, _f0, f1, _f1, f2, _f2;
  char buffer[256] = {0};
  while (!in.getline(buffer, 255).eof()) {
    // A line of 255 or more characters sets failbit without eofbit; the
    // reference then loops forever on the failed stream.
    if (in.fail()) return ORC_E_PARSE;
    if (buffer[0] == 'v' && (buffer[1] == ' ' || buffer[1] == 32)) {
      if (std::sscanf(buffer, "v %f %f %f", &x, &y, &z) == 3) {
        P.push_back(x);
        P.push_back(y);
        P.push_back(z);
      }
    } else if (buffer[0] == 'f' && (buffer[1] == ' ' || buffer[1] == 32)) {
      if (std::sscanf(buffer, "f %d//%d %d//%d %d//%d", &f0, &_f0, &f1, &_f1, &f2, &_f2) == 6 ||
          std::sscanf(buffer, "f %d/%d %d/%d %d/%d", &f0, &_f0, &f1, &_f1, &f2, &_f2) == 6 ||
          std::sscanf(buffer, "f %d %d %d", &f0, &f1, &f2) == 3) {
        // f - 1 wraps for INT_MIN (undefined in the reference's int
        // arithmetic; two's complement in its compiled code)
        F.push_back((int32_t)((uint32_t)f0 - 1u));
        F.push_back((int32_t)((uint32_t)f1 - 1u));
        F.push_back((int32_t)((uint32_t)f2 - 1u));
      }
    }
  }
  *n_points = P.size() / 3;
  *n_tris = F.size() / 3;
  *points = (float*)std::malloc(std::max<size_t>(1, P.size() * sizeof(float)));
  *tris = (int32_t*)std::malloc(std::max<size_t>(1, F.size() * sizeof(int32_t)));
  if (!*points || !*tris) return ORC_E_OOM;
  if (!P.empty()) std::memcpy(*points, P.data(), P.size() * sizeof(float));
  if (!F.empty()) std::memcpy(*tris, F.data(), F.size() * sizeof(int32_t));
  return ORC_OK;
}

void oracle_free_buf(void* p) { std::free(p); }

void* oracle_build(const float* points, uint64_t n_points, const int32_t* tris, uint64_t n_tris,
                   const int32_t* res, int32_t semantics, int32_t key_bits, int32_t n_threads,
                   int32_t* status) {
  Tree* T = new (std::nothrow) Tree();
  if (!T) {
    if (status) *status = ORC_E_OOM;
    return nullptr;
  }
  int st = ORC_OK;
  try {
    st = build_tree(points, n_points, tris, n_tris, res, semantics, key_bits, n_threads, *T);
  } catch (const std::bad_alloc&) {
    st = ORC_E_OOM;
  }
  T->info.status = st;
  if (status) *status = st;
  return T;
}

void oracle_tree_free(void* tree) { delete (Tree*)tree; }

int oracle_tree_info(const void* tree, oracle_info* out) {
  *out = ((const Tree*)tree)->info;
  return ORC_OK;
}

uint64_t oracle_tree_field_bytes(const void* tree, int32_t field) {
  const Tree& T = *(const Tree*)tree;
  const uint64_t N = T.code.size();
  const uint64_t kb = T.info.key_bits == 64 ? 8 : 4;
  switch (field) {
    case ORC_F_VOXELS: return kb * T.voxels.size();
    case ORC_F_CODE: return kb * N;
    case ORC_F_ISLEAF: return N;
    case ORC_F_ORIGIN: return 12 * N;
    case ORC_F_WIDTH: return 4 * N;
    case ORC_F_PARENT: return 4 * N;
    case ORC_F_CHILDS: return 32 * N;
    case ORC_F_NEIGHBORS: return 108 * N;
    case ORC_F_VERTICES: return 16 * T.verts.size();
    case ORC_F_EDGES: return 28 * T.edges.size();
    case ORC_F_FACES: return 28 * T.faces.size();
    case ORC_F_LEVEL_NODES: return 8 * T.level_nodes.size();
    case ORC_F_NODES_AOS: return 168 * N;
    default: return 0;
  }
}

int oracle_tree_get_field(const void* tree, int32_t field, void* dst, uint64_t dst_bytes) {
  const Tree& T = *(const Tree*)tree;
  const uint64_t need = oracle_tree_field_bytes(tree, field);
  if (dst_bytes < need) return ORC_E_INVALID;
  const bool k32 = T.info.key_bits != 64;
  auto put_codes = [&](const std::vector<uint64_t>& v) {
    if (k32) {
      uint32_t* d = (uint32_t*)dst;
      for (size_t i = 0; i < v.size(); ++i) d[i] = (uint32_t)v[i];
    } else if (!v.empty()) {
      std::memcpy(dst, v.data(), 8 * v.size());
    }
  };
  auto put = [&](const void* src) {
    if (need) std::memcpy(dst, src, need);
  };
  switch (field) {
    case ORC_F_VOXELS: put_codes(T.voxels); break;
    case ORC_F_CODE: put_codes(T.code); break;
    case ORC_F_ISLEAF: put(T.is_leaf.data()); break;
    case ORC_F_ORIGIN: put(T.origin.data()); break;
    case ORC_F_WIDTH: put(T.width.data()); break;
    case ORC_F_PARENT: put(T.parent.data()); break;
    case ORC_F_CHILDS: put(T.childs.data()); break;
    case ORC_F_NEIGHBORS: put(T.neighbors.data()); break;
    case ORC_F_VERTICES: put(T.verts.data()); break;
    case ORC_F_EDGES: put(T.edges.data()); break;
    case ORC_F_FACES: put(T.faces.data()); break;
    case ORC_F_LEVEL_NODES: put(T.level_nodes.data()); break;
    case ORC_F_NODES_AOS: {
      SVONodeRec* d = (SVONodeRec*)dst;
      for (uint64_t t = 0; t < T.code.size(); ++t) {
        SVONodeRec r;
        std::memset(&r, 0, sizeof(r));
        r.mortonCode = (uint32_t)T.code[t];
        r.isLeaf = T.is_leaf[t];
        for (int a = 0; a < 3; ++a) r.origin[a] = T.origin[3 * t + a];
        r.width = T.width[t];
        r.parent = T.parent[t];
        for (int c = 0; c < 8; ++c) r.childs[c] = T.childs[8 * t + c];
        for (int s = 0; s < 27; ++s) r.neighbors[s] = T.neighbors[27 * t + s];
        d[t] = r;
      }
      break;
    }
    default: return ORC_E_INVALID;
  }
  return ORC_OK;
}

int oracle_write_obj(const void* tree, const char* path) { return write_tree(*(const Tree*)tree, path); }
int oracle_write_voxel_obj(const void* tree, const char* path) {
  return write_voxels(*(const Tree*)tree, path);
}

uint32_t oracle_morton_encode_u16(uint32_t x, uint32_t y, uint32_t z) {
  return encode_u16((uint16_t)x, (uint16_t)y, (uint16_t)z);
}
uint64_t oracle_morton_encode_u64(uint32_t x, uint32_t y, uint32_t z) { return encode_u64(x, y, z); }
void oracle_morton_decode(uint64_t code, uint32_t* x, uint32_t* y, uint32_t* z) { decode(code, *x, *y, *z); }

void oracle_luts(uint32_t* enc_x256, uint32_t* enc_y256, uint32_t* enc_z256, uint8_t* dec_x512,
                 uint8_t* dec_y512, uint8_t* dec_z512, int16_t* nb_parent_8x27,
                 int16_t* nb_child_8x27, int16_t* vert_shared_8x8, int16_t* edge_shared_12x4) {
  const Luts& L = luts();
  std::memcpy(enc_x256, L.enc[0], sizeof(L.enc[0]));
  std::memcpy(enc_y256, L.enc[1], sizeof(L.enc[1]));
  std::memcpy(enc_z256, L.enc[2], sizeof(L.enc[2]));
  std::memcpy(dec_x512, L.dec[0], sizeof(L.dec[0]));
  std::memcpy(dec_y512, L.dec[1], sizeof(L.dec[1]));
  std::memcpy(dec_z512, L.dec[2], sizeof(L.dec[2]));
  std::memcpy(nb_parent_8x27, L.nb_parent, sizeof(L.nb_parent));
  std::memcpy(nb_child_8x27, L.nb_child, sizeof(L.nb_child));
  std::memcpy(vert_shared_8x8, L.vert_shared, sizeof(L.vert_shared));
  std::memcpy(edge_shared_12x4, L.edge_shared, sizeof(L.edge_shared));
}

void oracle_set_arith_variant(int32_t variant) { g_arith = variant & 3; }

uint64_t oracle_voxelize(const float* points, uint64_t n_points, const int32_t* tris,
                         uint64_t n_tris, const int32_t* res, int32_t key_bits, uint64_t* dst,
                         uint64_t cap, int32_t* status) {
  int st = validate(points, n_points, tris, n_tris, res, key_bits);
  Grid g;
  if (!st) st = make_grid(points, n_points, res, g);
  std::vector<uint64_t> v;
  if (!st) st = voxelize_all(points, tris, n_tris, g, key_bits, 1, v);
  if (status) *status = st;
  if (st && st != ORC_E_EMPTY) return 0;
  for (uint64_t i = 0; i < v.size() && i < cap; ++i) dst[i] = v[i];
  return v.size();
}

uint64_t* oracle_voxelize_alloc(const float* points, uint64_t n_points, const int32_t* tris,
                                uint64_t n_tris, const int32_t* res, int32_t key_bits,
                                int32_t n_threads, uint64_t* n_out, int32_t* status) {
  int st = validate(points, n_points, tris, n_tris, res, key_bits);
  Grid g;
  if (!st) st = make_grid(points, n_points, res, g);
  std::vector<uint64_t> v;
  if (!st) st = voxelize_all(points, tris, n_tris, g, key_bits, n_threads, v);
  if (status) *status = st;
  *n_out = 0;
  if (st && st != ORC_E_EMPTY) return nullptr;
  uint64_t* out = (uint64_t*)std::malloc(sizeof(uint64_t) * (v.size() ? v.size() : 1));
  if (!out) {
    if (status) *status = ORC_E_OOM;
    return nullptr;
  }
  std::memcpy(out, v.data(), sizeof(uint64_t) * v.size());
  *n_out = v.size();
  return out;
}

}  // extern "C"
