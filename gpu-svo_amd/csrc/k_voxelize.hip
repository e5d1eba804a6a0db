// k_voxelize.hip -- point bbox, grid parameters and surface voxelization.
//
// Follows the reference's surfaceVoxelize (SVO.cu:37-188): per triangle a
// plane test and three 2-D projection tests over every cell of the clamped
// bounding box, with the same float operation order (built with
// -ffp-contract=off, correctly rounded divide and sqrt, denormals kept).
//
// Work distribution (the reference gives one thread per triangle):
//  * k_classify   one lane per triangle: box cells -> key capacity and the
//                 number of big-triangle chunks (exact, read back once).
//  * k_vox_small  one lane per triangle whose box has <= small_max cells,
//                 the workgroup's triangles handed out in order of box size;
//                 lanes walk their box row by row, and the yz-projection
//                 test, which does not depend on x, is evaluated once per
//                 row (an exact hoist: the same operands, the same rounding).
//                 Larger triangles are cut into chunks of 512 rows and queued.
//  * k_vox_big    one wave64 per queued item of 512 rows: each lane takes
//                 eight rows, prunes each to the x-range outside of which a
//                 test provably fails (row_range) and tests the cells left
//                 in it.
// Hits are staged per wave in LDS and flushed with one atomic per ~1K keys.
#include "svo_internal.h"

namespace svo {
namespace {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int STAGE = 1024;        // keys staged per wave
constexpr int BBOX_CHUNK = 4096;   // points per bbox block

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t bcast_first(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint32_t find_mesh(const uint64_t* base, uint32_t n_meshes, uint64_t gid) {
  uint32_t lo = 0, hi = n_meshes;  // base[lo] <= gid < base[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= gid) lo = mid; else hi = mid;
  }
  return lo;
}

// Morton code stepping: the bits of one axis are a dilated integer, so adding
// one to that axis is an add through the other axes' bits (set to one so the
// carry passes them) and a merge back.  x is bits 0, 3, 6, ...; y and z the
// same shifted by one and two.
template <typename K>
__device__ inline K morton_x_mask() {
  return sizeof(K) == 4 ? (K)0x09249249u : (K)0x1249249249249249ull;
}
template <typename K>
__device__ inline K dilated_inc(K code, K mask) {
  return (((code | ~mask) + (K)1) & mask) | (code & ~mask);
}

// ------------------------------------------------------------------ bbox
__global__ void k_bbox_init(uint32_t* acc, uint32_t n_meshes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_meshes * 8) return;
  const uint32_t f = i & 7;
  acc[i] = f < 3 ? 0xFFFFFFFFu : 0u;  // ordered min starts at +max, max at 0
}

__global__ __launch_bounds__(BLOCK) void k_bbox(const MeshDev* meshes, uint32_t* acc) {
  const MeshDev md = meshes[blockIdx.y];
  const uint64_t start = (uint64_t)blockIdx.x * BBOX_CHUNK;
  if (start >= md.n_points) return;
  const uint64_t end = start + BBOX_CHUNK < md.n_points ? start + BBOX_CHUNK : md.n_points;
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0, 0, 0};
  uint32_t bad = 0;
  for (uint64_t i = start + threadIdx.x; i < end; i += BLOCK) {
    for (int a = 0; a < 3; ++a) {
      const float v = md.pts[3 * i + a];
      if (!isfinite(v)) {
        bad = 1;
        continue;
      }
      const uint32_t k = float_to_ordered(v);
      mn[a] = k < mn[a] ? k : mn[a];
      mx[a] = k > mx[a] ? k : mx[a];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int a = 0; a < 3; ++a) {
      const uint32_t omn = __shfl_xor(mn[a], off);
      const uint32_t omx = __shfl_xor(mx[a], off);
      mn[a] = omn < mn[a] ? omn : mn[a];
      mx[a] = omx > mx[a] ? omx : mx[a];
    }
    bad |= __shfl_xor(bad, off);
  }
  __shared__ uint32_t red[WAVES][7];
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    for (int a = 0; a < 3; ++a) {
      red[w][a] = mn[a];
      red[w][3 + a] = mx[a];
    }
    red[w][6] = bad;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    uint32_t v = red[0][threadIdx.x];
    for (int k = 1; k < WAVES; ++k) {
      const uint32_t o = red[k][threadIdx.x];
      if (threadIdx.x < 3) v = o < v ? o : v;
      else if (threadIdx.x < 6) v = o > v ? o : v;
      else v |= o;
    }
    uint32_t* a = acc + 8 * blockIdx.y;
    if (threadIdx.x < 3) atomicMin(a + threadIdx.x, v);
    else if (threadIdx.x < 6) atomicMax(a + threadIdx.x, v);
    else if (v) atomicOr(a + 6, 1u);
  }
}

// createMeshBBCube (BaseModel.h:24-49) and the voxel size (SVO.cu:348-350).
__global__ void k_params(const MeshDev* meshes, uint32_t n_meshes, const uint32_t* acc,
                         MeshParams* params) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_meshes) return;
  const uint32_t* a = acc + 8 * m;
  MeshParams P;
  P.status = a[6] ? ERR_NONFINITE : 0u;
  float mn[3], mx[3], len[3], amin[3], amax[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = ordered_to_float(a[k]);
    mx[k] = ordered_to_float(a[3 + k]);
    len[k] = mx[k] - mn[k];
    amin[k] = mn[k];
    amax[k] = mx[k];
    P.res[k] = meshes[m].res[k];
  }
  const float max_length = fmax_r(len[0], fmax_r(len[1], len[2]));
  if (!(max_length > 0.0f)) P.status |= ERR_DEGENERATE;
  for (int k = 0; k < 3; ++k) {
    if (max_length == len[k]) continue;
    const float delta = max_length - len[k];
    amin[k] = mn[k] - (delta / 2.0f);
    amax[k] = mx[k] + (delta / 2.0f);
  }
  for (int k = 0; k < 3; ++k) {
    const float eps = (amax[k] - amin[k]) / 10001.0f;
    amin[k] -= eps;
    amax[k] += eps;
  }
  for (int k = 0; k < 3; ++k) {
    P.bmin[k] = amin[k];
    P.bmax[k] = amax[k];
    P.width[k] = amax[k] - amin[k];
    P.origin[k] = amin[k];
    P.unit[k] = P.width[k] / (float)P.res[k];
  }
  params[m] = P;
}

// ------------------------------------------------------------------ triangle setup
struct TriSetup {
  float n[3];
  float d1, d2;
  float ne[3][3][2];  // [plane xy|yz|zx][edge][2]
  float de[3][3];
  float u[3];
  int g0[3], g1[3];
};

// Loads a triangle relative to gridOrigin (SVO.cu:58-63).  false on a bad index.
__device__ inline bool load_tri(const MeshDev& md, const MeshParams& P, uint64_t local,
                                float v[3][3]) {
  const int32_t* t = md.tri + 3 * local;
  for (int k = 0; k < 3; ++k) {
    const int32_t i = t[k];
    if (i < 0 || (uint64_t)i >= md.n_points) return false;
    for (int a = 0; a < 3; ++a) v[k][a] = md.pts[3 * (uint64_t)i + a] - P.origin[a];
  }
  return true;
}

// Triangle box in grid cells (SVO.cu:73-85).
__device__ inline void tri_box(const float v[3][3], const MeshParams& P, int g0[3], int g1[3]) {
  for (int a = 0; a < 3; ++a) {
    const float lo = fmin_r(v[0][a], fmin_r(v[1][a], v[2][a]));
    const float hi = fmax_r(v[0][a], fmax_r(v[1][a], v[2][a]));
    g0[a] = clamp_i((int)(lo / P.unit[a]), 0, P.res[a] - 1);
    g1[a] = clamp_i((int)(hi / P.unit[a]), 0, P.res[a] - 1);
  }
}

// Plane and projection test constants (SVO.cu:65-149).
__device__ inline void tri_setup(const float v[3][3], const MeshParams& P, TriSetup& S) {
  float e[3][3];
  for (int a = 0; a < 3; ++a) {
    e[0][a] = v[1][a] - v[0][a];
    e[1][a] = v[2][a] - v[1][a];
    e[2][a] = v[0][a] - v[2][a];
    S.u[a] = P.unit[a];
  }
  float n0 = e[0][1] * e[1][2] - e[0][2] * e[1][1];
  float n1 = e[0][2] * e[1][0] - e[0][0] * e[1][2];
  float n2 = e[0][0] * e[1][1] - e[0][1] * e[1][0];
  const float zz = n0 * n0 + (n1 * n1 + n2 * n2);  // squaredNorm, a0 + (a1 + a2)
  if (zz > 0.0f) {
    const float s = sqrtf(zz);
    n0 = n0 / s;
    n1 = n1 / s;
    n2 = n2 / s;
  }
  S.n[0] = n0;
  S.n[1] = n1;
  S.n[2] = n2;
  tri_box(v, P, S.g0, S.g1);
  const float c0 = n0 > 0.0f ? S.u[0] : 0.0f;
  const float c1 = n1 > 0.0f ? S.u[1] : 0.0f;
  const float c2 = n2 > 0.0f ? S.u[2] : 0.0f;
  S.d1 = n0 * (c0 - v[0][0]) + (n1 * (c1 - v[0][1]) + n2 * (c2 - v[0][2]));
  S.d2 = n0 * ((S.u[0] - c0) - v[0][0]) +
         (n1 * ((S.u[1] - c1) - v[0][1]) + n2 * ((S.u[2] - c2) - v[0][2]));
  const bool flip[3] = {n2 < 0.0f, n0 < 0.0f, n1 < 0.0f};
  // plane p uses axes (a, b): xy = (0, 1), yz = (1, 2), zx = (2, 0)
  for (int p = 0; p < 3; ++p) {
    const int a = p, b = (p + 1) % 3;
    for (int i = 0; i < 3; ++i) {
      float x0 = -1.0f * e[i][b];
      float x1 = e[i][a];
      if (flip[p]) {
        x0 = -x0;
        x1 = -x1;
      }
      S.ne[p][i][0] = x0;
      S.ne[p][i][1] = x1;
      S.de[p][i] = ((-1.0f * (x0 * v[i][a] + x1 * v[i][b])) + fmax_r(0.0f, S.u[a] * x0)) +
                   fmax_r(0.0f, S.u[b] * x1);
    }
  }
}

// Row terms of cell row (y, z): everything of the 13 tests that does not
// depend on x.  Returns false when a yz-projection test fails (the whole row
// is empty).
struct RowTerms {
  float A;       // n.y*p.y + n.z*p.z  (inner sum of n.dot(p))
  float rxy[3];  // ne_xy[i].y * p.y
  float rzx[3];  // ne_zx[i].x * p.z
};
// The tests are combined without short-circuit: every lane evaluates all of
// them (the same operations and roundings), so the wave runs no branches.
__device__ inline bool row_terms(const TriSetup& S, int y, int z, RowTerms& R) {
  const float py = (float)y * S.u[1];
  const float pz = (float)z * S.u[2];
  bool ok = true;
  for (int i = 0; i < 3; ++i)
    ok &= !((S.ne[1][i][0] * py + S.ne[1][i][1] * pz) + S.de[1][i] < 0.0f);
  R.A = S.n[1] * py + S.n[2] * pz;
  for (int i = 0; i < 3; ++i) {
    R.rxy[i] = S.ne[0][i][1] * py;
    R.rzx[i] = S.ne[2][i][0] * pz;
  }
  return ok;
}
__device__ inline bool cell_hit(const TriSetup& S, const RowTerms& R, int x) {
  const float px = (float)x * S.u[0];
  const float t = S.n[0] * px + R.A;
  bool ok = !((t + S.d1) * (t + S.d2) > 0.0f);
  for (int i = 0; i < 3; ++i) {
    ok &= !((S.ne[0][i][0] * px + R.rxy[i]) + S.de[0][i] < 0.0f);
    ok &= !((R.rzx[i] + S.ne[2][i][1] * px) + S.de[2][i] < 0.0f);
  }
  return ok;
}

// Conservative x-range [xa, xb] of row (y, z): outside it the plane test or
// one of the xy / zx edge tests fails for certain.  Each test is affine in x;
// its root is computed in double and widened by two cells plus a bound on
// the float rounding of the test (2^-20 of the magnitudes involved, several
// times the few half-ulp roundings of the float evaluation).  Tests whose
// x-slope is tiny are not used.  Skipping cells outside the range therefore
// never changes the set of hits.  The x-slopes and their reciprocals depend
// on the triangle only (RangeConst, once per triangle); a root is the row
// term times the reciprocal, within a few double ulps of the quotient, far
// inside the margin.
struct RangeConst {
  double a, ia;        // plane test: slope n.x * u.x and its reciprocal
  double b[6], ib[6];  // xy edges 0-2, zx edges 0-2
};

__device__ inline void range_const(const TriSetup& S, RangeConst& C) {
  const double ux = S.u[0];
  C.a = (double)S.n[0] * ux;
  C.ia = C.a != 0.0 ? 1.0 / C.a : 0.0;
  for (int i = 0; i < 6; ++i) {
    C.b[i] = (double)(i < 3 ? S.ne[0][i][0] : S.ne[2][i - 3][1]) * ux;
    C.ib[i] = C.b[i] != 0.0 ? 1.0 / C.b[i] : 0.0;
  }
}

__device__ inline void row_range(const TriSetup& S, const RangeConst& C, const RowTerms& R,
                                 int& xa, int& xb) {
  double lo = S.g0[0], hi = S.g1[0];
  const double xmax = (double)S.g1[0] + 1.0;
  {
    const double a = C.a;
    const double d1 = S.d1, d2 = S.d2, A = R.A;
    const double mag = fabs(a) * xmax + fabs(A) + fabs(d1) + fabs(d2);
    if (fabs(a) > 1e-12 * (mag + 1e-30) && fabs(a) > 1e-15) {
      const double err = mag * 0x1p-20 * fabs(C.ia) + 2.0;
      double x1 = (-fmax(d1, d2) - A) * C.ia, x2 = (-fmin(d1, d2) - A) * C.ia;
      if (x1 > x2) {
        const double t = x1;
        x1 = x2;
        x2 = t;
      }
      lo = fmax(lo, floor(x1 - err));
      hi = fmin(hi, ceil(x2 + err));
    }
  }
  for (int i = 0; i < 6; ++i) {
    // xy: ne.x * px + rxy + de >= 0;  zx: rzx + ne.y * px + de >= 0
    const double b = C.b[i];
    const double c = i < 3 ? (double)R.rxy[i] + (double)S.de[0][i]
                           : (double)R.rzx[i - 3] + (double)S.de[2][i - 3];
    const double mag = fabs(b) * xmax + fabs(c) + fabs(i < 3 ? R.rxy[i] : R.rzx[i - 3]);
    if (!(fabs(b) > 1e-12 * (mag + 1e-30) && fabs(b) > 1e-30)) continue;
    const double err = mag * 0x1p-20 * fabs(C.ib[i]) + 2.0;
    const double root = -c * C.ib[i];
    if (b > 0) lo = fmax(lo, floor(root - err));
    else hi = fmin(hi, ceil(root + err));
  }
  xa = lo > (double)S.g1[0] ? S.g1[0] + 1 : (int)lo;
  xb = hi < (double)S.g0[0] ? S.g0[0] - 1 : (int)hi;
}

template <typename K>
struct Stager {
  K* buf;            // this wave's LDS stage
  uint32_t n = 0;    // wave-uniform count
  K* keys;
  unsigned long long* count;
  __device__ inline void flush() {
    wave_sync();
    uint64_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(count, (unsigned long long)n);
    base = bcast_first(base);
    for (uint32_t i = __lane_id(); i < n; i += 64) keys[base + i] = buf[i];
    wave_sync();
    n = 0;
  }
  __device__ inline void push(bool hit, K key) {
    const uint64_t mask = __ballot(hit);
    if (mask == 0) return;
    if (hit) buf[n + lanes_below(mask)] = key;
    n += (uint32_t)__popcll(mask);
    if (n > STAGE - 64) flush();
  }
  __device__ inline void finish() {
    if (n) flush();
  }
};

// Occupancy-bitmap output: bit (code & 31) of word bm_off[mesh] + code / 32
// is set for every hit (the dense flag array of SVO.cu:180-182 at one bit per
// cell).  Each lane keeps the word of its last hit and ORs it into memory
// only when a hit falls in another word (consecutive cells of a row share a
// word up to four at a time), so the atomics issued are fewer than the hits.
// Raw hits are counted per wave and added once at the end.
struct BitSink {
  uint32_t* bits;
  unsigned long long* count;
  uint64_t n = 0;          // wave-uniform
  uint64_t pend_word = 0;  // per lane
  uint32_t pend_bits = 0;
  __device__ inline void flush_lane() {
    if (pend_bits)
      __hip_atomic_fetch_or(bits + pend_word, pend_bits, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ inline void push(bool hit, uint64_t word, uint32_t bit) {
    const uint64_t mask = __ballot(hit);
    if (mask == 0) return;
    n += (uint64_t)__popcll(mask);
    if (hit) {
      if (word != pend_word) {
        flush_lane();
        pend_word = word;
        pend_bits = 0;
      }
      pend_bits |= 1u << bit;
    }
  }
  __device__ inline void finish() {
    flush_lane();
    pend_bits = 0;
    if (n && __lane_id() == 0) atomicAdd(count, (unsigned long long)n);
  }
};

// Where a voxelization kernel puts its hits: keys (staged, for the sort) or
// the occupancy bitmap.
template <typename K>
struct VoxOut {
  K* keys;                // key path
  uint32_t* bits;         // bitmap path
  const uint64_t* bm_off; // bitmap words before each mesh
  unsigned long long* count;
};

// Big-triangle work item: 512 rows (y, z) of the box, eight per lane (256
// and 128 rows measured slower: more items, each paying the triangle set-up).
constexpr uint64_t BIG_ROWS = 512;

// Voxelization order.  One lane walks each small triangle's box, so a wave
// runs as long as the largest box among its 64 triangles.  k_classify gives
// every triangle a size class (one byte: triangles past the one-lane limit,
// which k_vox_small only queues, first, then eighth-octave classes of the box
// cells, exact below 32 cells, largest first), and k_order sorts each run of
// ORDER_GROUP consecutive triangles by class into `order`: the waves of a run
// see boxes of about one size.  Triangles of one class keep roughly their
// input order, which keeps the neighbouring triangles of a mesh together.  k_vox_small deals the runs' workgroup slots
// out slot-major (the first slot of every run, then the second, ...), so the
// workgroups of the largest boxes start first and the launch ends on the
// smallest.  The order does not change the hits.
constexpr uint32_t ORDER_THREADS = 1024;
constexpr uint32_t ORDER_PER = 8;  // triangles per thread of k_order
constexpr uint32_t ORDER_GROUP = ORDER_THREADS * ORDER_PER;
constexpr uint32_t ORDER_BINS = 256;
static_assert(ORDER_GROUP % BLOCK == 0, "a run fills whole workgroups of k_vox_small");

// Size class, largest boxes first: 0 past the one-lane limit, then
// 254 - (eighth-octave class of the cells); 255 for triangles that walk
// nothing (failed mesh, bad index).
__device__ inline uint32_t size_bin(uint64_t cells, uint64_t small_max) {
  if (cells > small_max) return 0;
  if (cells < 32) return ORDER_BINS - 2 - (uint32_t)cells;
  const uint32_t e = 63u - (uint32_t)__builtin_clzll(cells);  // >= 5
  const uint32_t b = 32u + 8u * (e - 5u) + (uint32_t)((cells >> (e - 3u)) & 7u);
  return ORDER_BINS - 2 - (b < ORDER_BINS - 3 ? b : ORDER_BINS - 3);
}

// A pass of at most 8 runs (65,536 triangles) keeps the triangles' own
// order: there the run sort's kernel (7-9 us, on the pass's critical path)
// costs more than the voxelizer saves (256 single-mesh passes of 320-81,920
// triangles took 80 ms sorted against 71-79 ms unsorted).
inline bool sorted_order(uint64_t n_tris) { return n_tris > 8ull * ORDER_GROUP; }

// Candidate cells and big-triangle chunk counts (added to totals[0],
// totals[1]) and each triangle's size class (bins[gid], when bins is set).
// A grid-stride loop over a bounded grid: the two totals take one atomic per
// workgroup, and atomics on one address serialize.
constexpr uint32_t CLASSIFY_BLOCKS = 1024;
__global__ __launch_bounds__(BLOCK) void k_classify(const MeshDev* meshes,
                                                    const MeshParams* params,
                                                    const uint64_t* tri_base, uint32_t n_meshes,
                                                    uint64_t n_tris, uint32_t small_max,
                                                    unsigned long long* totals, uint8_t* bins) {
  uint64_t cells = 0, chunks = 0;
  for (uint64_t gid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; gid < n_tris;
       gid += (uint64_t)gridDim.x * BLOCK) {
    uint32_t bin = ORDER_BINS - 1;
    const uint32_t m = find_mesh(tri_base, n_meshes, gid);
    const MeshDev& md = meshes[m];
    const MeshParams& P = params[m];
    float v[3][3];
    if (P.status == 0 && load_tri(md, P, gid - md.tri_base, v)) {
      int g0[3], g1[3];
      tri_box(v, P, g0, g1);
      const uint64_t nx = g1[0] - g0[0] + 1, ny = g1[1] - g0[1] + 1, nz = g1[2] - g0[2] + 1;
      const uint64_t c = nx * ny * nz;
      const uint64_t sm = md.small_max ? md.small_max : small_max;
      cells += c;
      if (c > sm) chunks += (ny * nz + BIG_ROWS - 1) / BIG_ROWS;
      bin = size_bin(c, sm);
    }
    if (bins) bins[gid] = (uint8_t)bin;
  }
  for (int off = 32; off > 0; off >>= 1) {
    cells += __shfl_xor(cells, off);
    chunks += __shfl_xor(chunks, off);
  }
  __shared__ uint64_t red[WAVES][2];
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6][0] = cells;
    red[threadIdx.x >> 6][1] = chunks;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint64_t t = 0;
    for (int w = 0; w < WAVES; ++w) t += red[w][threadIdx.x];
    if (t) atomicAdd(totals + threadIdx.x, (unsigned long long)t);
  }
}

// One workgroup per run of ORDER_GROUP triangles: a counting sort of the
// run's size classes into `order`.  Each wave counts its triangles in a
// histogram of its own (a run's triangles fall into few classes, and LDS
// atomics on one address serialize across the whole workgroup), the
// (class, wave) counts are scanned class-major, and every triangle lands at
// its class's and wave's offset plus its rank in the wave's count.  Thread t
// takes triangles t, t + 1024, ... of the run, so the lanes of a wave hold
// consecutive triangles and a class's triangles from one wave land near each
// other (the order of ranks within one atomic instruction is the hardware's).
constexpr uint32_t ORDER_WAVES = ORDER_THREADS / 64;
__global__ __launch_bounds__(ORDER_THREADS) void k_order(const uint8_t* bins, uint64_t n_tris,
                                                         uint32_t* order) {
  __shared__ uint32_t s_cnt[ORDER_BINS][ORDER_WAVES];  // then (class, wave) offsets
  __shared__ uint32_t s_tot[ORDER_BINS];
  const uint32_t wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < ORDER_BINS * ORDER_WAVES; i += ORDER_THREADS)
    (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t g0 = (uint64_t)blockIdx.x * ORDER_GROUP;
  uint32_t bin[ORDER_PER], rank[ORDER_PER];
#pragma unroll
  for (uint32_t k = 0; k < ORDER_PER; ++k) {
    const uint64_t gid = g0 + k * ORDER_THREADS + threadIdx.x;
    bin[k] = gid < n_tris ? bins[gid] : 0u;
  }
#pragma unroll
  for (uint32_t k = 0; k < ORDER_PER; ++k)
    rank[k] = g0 + k * ORDER_THREADS + threadIdx.x < n_tris
                  ? atomicAdd(&s_cnt[bin[k]][wave], 1u) : 0u;
  __syncthreads();
  // per class: exclusive prefix over the waves, total to s_tot
  if (threadIdx.x < ORDER_BINS) {
    uint32_t t = 0;
#pragma unroll
    for (uint32_t w = 0; w < ORDER_WAVES; ++w) {
      const uint32_t c = s_cnt[threadIdx.x][w];
      s_cnt[threadIdx.x][w] = t;
      t += c;
    }
    s_tot[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    // exclusive scan of the class totals, four classes per lane
    const uint32_t l = threadIdx.x;
    uint32_t c[4], t = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      c[j] = s_tot[4 * l + j];
      t += c[j];
    }
    uint32_t inc = t;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(inc, d);
      if (l >= (uint32_t)d) inc += o;
    }
    uint32_t ex = inc - t;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      s_tot[4 * l + j] = ex;
      ex += c[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < ORDER_PER; ++k) {
    const uint64_t gid = g0 + k * ORDER_THREADS + threadIdx.x;
    if (gid < n_tris) order[g0 + s_tot[bin[k]] + s_cnt[bin[k]][wave] + rank[k]] = (uint32_t)gid;
  }
}

__global__ void k_flag_bad_index(const MeshDev* meshes, MeshParams* params,
                                 const uint64_t* tri_base, uint32_t n_meshes, uint64_t n_tris) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n_tris) return;
  const uint32_t m = find_mesh(tri_base, n_meshes, gid);
  const MeshDev& md = meshes[m];
  const int32_t* t = md.tri + 3 * (gid - md.tri_base);
  for (int k = 0; k < 3; ++k)
    if (t[k] < 0 || (uint64_t)t[k] >= md.n_points) atomicOr(&params[m].status, ERR_INDEX);
}

template <typename K, bool BITMAP>
__global__ __launch_bounds__(BLOCK) void k_vox_small(
    const MeshDev* meshes, const MeshParams* params, const uint64_t* tri_base, uint32_t n_meshes,
    uint64_t n_tris, int cb, uint32_t small_max, const uint32_t* order, VoxOut<K> out,
    BigItem* big, unsigned long long* big_count, uint64_t big_cap) {
  __shared__ K stage[BITMAP ? 1 : WAVES][BITMAP ? 1 : STAGE];
  Stager<K> st;
  st.buf = stage[BITMAP ? 0 : threadIdx.x >> 6];
  st.keys = out.keys;
  st.count = out.count;
  BitSink bs;
  bs.bits = out.bits;
  bs.count = out.count;
  // triangles in the order k_classify wrote (boxes of about one size per
  // wave), workgroup slots of the runs dealt out slot-major
  const uint32_t runs = (uint32_t)((n_tris + ORDER_GROUP - 1) / ORDER_GROUP);
  const uint64_t pos = (uint64_t)(blockIdx.x % runs) * ORDER_GROUP +
                       (uint64_t)(blockIdx.x / runs) * BLOCK + threadIdx.x;
  const uint64_t gid = pos >= n_tris ? n_tris : order ? order[pos] : pos;
  TriSetup S;
  uint64_t count = 0;
  K mesh_key = 0;
  uint64_t wbase = 0;
  if (gid < n_tris) {
    const uint32_t m = find_mesh(tri_base, n_meshes, gid);
    const MeshDev& md = meshes[m];
    const MeshParams P = params[m];
    float v[3][3];
    if (P.status == 0 && load_tri(md, P, gid - md.tri_base, v)) {
      tri_setup(v, P, S);
      const uint64_t nx = S.g1[0] - S.g0[0] + 1, ny = S.g1[1] - S.g0[1] + 1,
                     nz = S.g1[2] - S.g0[2] + 1;
      const uint64_t cells = nx * ny * nz;
      if (cells <= (md.small_max ? md.small_max : small_max)) {
        count = cells;
      } else {
        const uint64_t nch = (ny * nz + BIG_ROWS - 1) / BIG_ROWS;
        const uint64_t at = atomicAdd(big_count, (unsigned long long)nch);
        for (uint64_t j = 0; j < nch; ++j)
          if (at + j < big_cap) big[at + j] = BigItem{(uint32_t)gid, (uint32_t)(j * BIG_ROWS)};
      }
      mesh_key = (K)m << cb;
      if (BITMAP) wbase = out.bm_off[m];
    }
  }
  uint64_t c = 0;
  int x = S.g0[0], y = S.g0[1], z = S.g0[2];
  bool row_ok = false;
  RowTerms R;
  // the Morton code of (x, y, z) as an x part (back to mx0 at each row start)
  // and a y|z part stepped row by row
  const K MX = morton_x_mask<K>(), MY = MX << 1, MZ = MX << 2;
  K mx0 = 0, my0 = 0, cyz = 0;
  if (count) {
    mx0 = morton3<K>((uint32_t)S.g0[0], 0u, 0u);
    my0 = morton3<K>(0u, (uint32_t)S.g0[1], 0u);
    cyz = my0 | morton3<K>(0u, 0u, (uint32_t)S.g0[2]);
  }
  K cx = mx0;
  while (__any(c < count)) {
    bool hit = false;
    K code = 0;
    if (c < count) {
      if (x == S.g0[0]) row_ok = row_terms(S, y, z, R);
      if (row_ok) {
        hit = cell_hit(S, R, x);
        code = cx | cyz;
      }
      if (row_ok && x < S.g1[0]) {
        ++x;
        ++c;
        cx = dilated_inc(cx, MX);
      } else {  // end of row, or an empty row skipped as a whole
        c += (uint64_t)(S.g1[0] - x + 1);
        x = S.g0[0];
        cx = mx0;
        if (++y > S.g1[1]) {
          y = S.g0[1];
          ++z;
          cyz = dilated_inc((cyz & ~MY) | my0, MZ);
        } else {
          cyz = dilated_inc(cyz, MY);
        }
      }
    }
    if (BITMAP) bs.push(hit, wbase + (uint64_t)(code >> 5), (uint32_t)(code & 31));
    else st.push(hit, mesh_key | code);
  }
  if (BITMAP) bs.finish();
  else st.finish();
}

// One wave per work item of 512 rows: each lane prunes eight rows of the item
// to their x-ranges and walks them; each inner loop runs as long as the
// longest pruned row of the 64.
template <typename K, bool BITMAP>
__global__ __launch_bounds__(BLOCK) void k_vox_big(const MeshDev* meshes,
                                                   const MeshParams* params,
                                                   const uint64_t* tri_base, uint32_t n_meshes,
                                                   int cb, const BigItem* big, uint64_t big_cap,
                                                   const unsigned long long* big_count,
                                                   uint64_t n_tris, VoxOut<K> out) {
  __shared__ K stage[BITMAP ? 1 : WAVES][BITMAP ? 1 : STAGE];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = __lane_id();
  Stager<K> st;
  st.buf = stage[BITMAP ? 0 : w];
  st.keys = out.keys;
  st.count = out.count;
  BitSink bs;
  bs.bits = out.bits;
  bs.count = out.count;
  const uint64_t item_id = (uint64_t)blockIdx.x * WAVES + w;
  const uint64_t queued = *big_count;
  const uint64_t n_items = queued < big_cap ? queued : big_cap;
  if (item_id < n_items && big[item_id].tri < n_tris) {
    const BigItem it = big[item_id];
    const uint32_t m = find_mesh(tri_base, n_meshes, it.tri);
    const MeshDev& md = meshes[m];
    const MeshParams P = params[m];
    float v[3][3];
    load_tri(md, P, it.tri - md.tri_base, v);  // indices were checked when queued
    TriSetup S;
    tri_setup(v, P, S);
    RangeConst RC;
    range_const(S, RC);
    const uint64_t ny = S.g1[1] - S.g0[1] + 1, nz = S.g1[2] - S.g0[2] + 1;
    const K mesh_key = (K)m << cb;
    const uint64_t wbase = BITMAP ? out.bm_off[m] : 0;
    for (uint64_t k = 0; k < BIG_ROWS; k += 64) {
      const uint64_t r = (uint64_t)it.row0 + k + lane;
      int y = 0, z = 0, xa = 0, xb = -1;
      RowTerms R;
      if (r < ny * nz) {
        y = S.g0[1] + (int)(r % ny);
        z = S.g0[2] + (int)(r / ny);
        if (row_terms(S, y, z, R)) row_range(S, RC, R, xa, xb);
      }
      // the code steps with x; past xb it is not used
      K code = morton3<K>((uint32_t)xa, (uint32_t)y, (uint32_t)z);
      const K MX = morton_x_mask<K>();
      for (int x = xa; __any(x <= xb); ++x) {
        const bool hit = (x <= xb) & cell_hit(S, R, x);
        if (BITMAP) bs.push(hit, wbase + (uint64_t)(code >> 5), (uint32_t)(code & 31));
        else st.push(hit, mesh_key | code);
        code = dilated_inc(code, MX);
      }
    }
  }
  if (BITMAP) bs.finish();
  else st.finish();
}

}  // namespace

void launch_bbox(const MeshDev* meshes, uint32_t n_meshes, uint64_t max_points, uint32_t* acc,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_bbox_init, dim3((n_meshes * 8 + 255) / 256), dim3(256), 0, s, acc,
                     n_meshes);
  if (max_points == 0) return;
  const uint32_t gx = (uint32_t)((max_points + BBOX_CHUNK - 1) / BBOX_CHUNK);
  hipLaunchKernelGGL(k_bbox, dim3(gx, n_meshes), dim3(BLOCK), 0, s, meshes, acc);
}

void launch_params(const MeshDev* meshes, uint32_t n_meshes, const uint32_t* acc,
                   MeshParams* params, hipStream_t s) {
  hipLaunchKernelGGL(k_params, dim3((n_meshes + 63) / 64), dim3(64), 0, s, meshes, n_meshes, acc,
                     params);
}

void launch_classify(const MeshDev* meshes, const MeshParams* params, const uint64_t* tri_base,
                     uint32_t n_meshes, uint64_t n_tris, const Tuning& tn,
                     unsigned long long* totals, uint32_t* order, uint8_t* bins, hipStream_t s) {
  if (n_tris == 0) return;
  const uint32_t blocks = (uint32_t)((n_tris + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_flag_bad_index, dim3(blocks), dim3(BLOCK), 0, s, meshes,
                     const_cast<MeshParams*>(params), tri_base, n_meshes, n_tris);
  hipLaunchKernelGGL(k_classify, dim3(std::min<uint32_t>(blocks, CLASSIFY_BLOCKS)), dim3(BLOCK), 0,
                     s, meshes, params, tri_base, n_meshes, n_tris, tn.small_max_cells, totals,
                     sorted_order(n_tris) ? bins : nullptr);
  if (!sorted_order(n_tris)) return;
  const uint32_t runs = (uint32_t)((n_tris + ORDER_GROUP - 1) / ORDER_GROUP);
  hipLaunchKernelGGL(k_order, dim3(runs), dim3(ORDER_THREADS), 0, s, (const uint8_t*)bins, n_tris,
                     order);
}

size_t order_bytes(uint64_t n_tris) {
  // order (uint32 per triangle), then the size classes (one byte each)
  const uint64_t n = std::max<uint64_t>(n_tris, 1);
  return ((4 * n + 255) & ~(uint64_t)255) + ((n + 255) & ~(uint64_t)255);
}

template <typename K>
void launch_voxelize(const MeshDev* meshes, const MeshParams* params, const uint64_t* tri_base,
                     uint32_t n_meshes, uint64_t n_tris, int cb, const Tuning& tn, K* keys,
                     uint32_t* bitmap, const uint64_t* bm_off, unsigned long long* key_count,
                     const uint32_t* order, BigItem* big, unsigned long long* big_count,
                     uint64_t big_cap, hipStream_t s) {
  if (n_tris == 0) return;
  const VoxOut<K> out{keys, bitmap, bm_off, key_count};
  if (!sorted_order(n_tris)) order = nullptr;  // the triangles' own order
  // every run of ORDER_GROUP triangles has ORDER_GROUP / BLOCK workgroup slots
  const uint32_t blocks = (uint32_t)((n_tris + ORDER_GROUP - 1) / ORDER_GROUP) *
                          (ORDER_GROUP / BLOCK);
  const uint32_t bblocks = (uint32_t)((big_cap + WAVES - 1) / WAVES);
  if (bitmap) {
    hipLaunchKernelGGL((k_vox_small<K, true>), dim3(blocks), dim3(BLOCK), 0, s, meshes, params,
                       tri_base, n_meshes, n_tris, cb, tn.small_max_cells, order, out,
                       big, big_count, big_cap);
    if (big_cap)
      hipLaunchKernelGGL((k_vox_big<K, true>), dim3(bblocks), dim3(BLOCK), 0, s, meshes, params,
                         tri_base, n_meshes, cb, (const BigItem*)big, big_cap,
                         (const unsigned long long*)big_count, n_tris, out);
  } else {
    hipLaunchKernelGGL((k_vox_small<K, false>), dim3(blocks), dim3(BLOCK), 0, s, meshes, params,
                       tri_base, n_meshes, n_tris, cb, tn.small_max_cells, order, out,
                       big, big_count, big_cap);
    if (big_cap)
      hipLaunchKernelGGL((k_vox_big<K, false>), dim3(bblocks), dim3(BLOCK), 0, s, meshes, params,
                         tri_base, n_meshes, cb, (const BigItem*)big, big_cap,
                         (const unsigned long long*)big_count, n_tris, out);
  }
}

template void launch_voxelize<uint32_t>(const MeshDev*, const MeshParams*, const uint64_t*,
                                        uint32_t, uint64_t, int, const Tuning&, uint32_t*,
                                        uint32_t*, const uint64_t*, unsigned long long*,
                                        const uint32_t*, BigItem*, unsigned long long*, uint64_t,
                                        hipStream_t);
template void launch_voxelize<uint64_t>(const MeshDev*, const MeshParams*, const uint64_t*,
                                        uint32_t, uint64_t, int, const Tuning&, uint64_t*,
                                        uint32_t*, const uint64_t*, unsigned long long*,
                                        const uint32_t*, BigItem*, unsigned long long*, uint64_t,
                                        hipStream_t);

}  // namespace svo
