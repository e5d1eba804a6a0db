/*
 * svo_kernels.h -- kernel-level entry points of libsvo_amd.so, for tests and
 * for callers that bring their own keys.  Host buffers in and out; the
 * library stages them through device memory and runs the same gfx950
 * kernels the build uses.
 *
 *   svo_k_radix_sort    LSD radix sort of Morton keys; in the reference the
 *                       sorted unique codes come from a dense flag array +
 *                       thrust::exclusive_scan + compactArray
 *                       (SVO.cu:367-405, 208-217).
 *   svo_k_voxelize      surface voxelization of one mesh (surfaceVoxelize,
 *                       SVO.cu:37-188): the raw hit codes.
 *   svo_k_ref_vertex_edge  vertex and edge records of nodes, reference
 *                       semantics (determineNodeVertex, determineNodeEdge,
 *                       SVO.cu:567-677).
 *   svo_k_octree_level  one level step on sorted codes: the unique codes
 *                       (scanMortonFlag + exclusive_scan + compactArray,
 *                       SVO.cu:30-35, 386-391, 208-217) and their parent
 *                       groups (getParentMorton + cpNumNodes, SVO.cu:22-24,
 *                       221-235, 426-437).
 */
#ifndef SVO_SVO_KERNELS_H
#define SVO_SVO_KERNELS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sorts n keys (key_bytes 4 or 8) on bits [0, bits) in place. */
int svo_k_radix_sort(int device, void* keys, uint64_t n, int key_bytes, int bits);

/* Sorts n device-resident keys (key_bytes 4 or 8) on bits [0, bits) in place
 * with the kernels of svo_k_radix_sort, allocating nothing: the key count,
 * the ping-pong buffer and the digit counts live in a fixed scratch array of
 * 34 MiB that each device holds from its first call until the process ends.
 * No counterpart in the reference; for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops).  n is at most
 * 4,194,304 (2^22) keys; larger n returns SVO_E_TOO_LARGE.  `stream` is a
 * hipStream_t of `device`: the call only enqueues on it, or, when it is
 * NULL, runs on the device's null stream and returns when the keys are
 * sorted.  All calls on one device share its scratch, so calls whose work
 * overlaps on the device (other streams, other threads) race.  The caller's
 * current device is restored on return. */
int svo_k_radix_sort_noalloc(int device, void* keys, uint64_t n, int key_bytes, int bits,
                             void* stream);

/* One octree level step on n sorted codes (key_bytes 4 or 8, duplicates
 * allowed, every code below 2^code_bits, 3 <= code_bits < 8 * key_bytes):
 * writes the n_unique distinct codes in order to `unique`, and for each of
 * the n_parents parent groups (codes 8p..8p+7 share parent p = code >> 3)
 * the parent code to `parents` and the index in `unique` of the group's
 * first member to `first_child`.  Each output holds up to n entries.
 * Unsorted input or a code of 2^code_bits or more is SVO_E_INVALID.  The
 * parents of one level are the sorted codes of the next: calling it again
 * on `parents` builds the level above. */
int svo_k_octree_level(int device, const void* sorted, uint64_t n, int key_bytes, int code_bits,
                       void* unique, uint64_t* n_unique, void* parents, uint32_t* first_child,
                       uint64_t* n_parents);

/* Vertex and edge records of n_nodes nodes as the reference builds them
 * (SVO.cu:567-677): with the reference's neighbor table (every slot empty
 * but the root's own) the owner of every record is its node.  Node i (origin
 * 3 floats, width 1 float) gets 8 vertex records at vertices[8i..8i+7]
 * (16 bytes each, svo_vertex in svo.h: the corner offset
 * w * (j & 1, j >> 1 & 1, j >> 2 & 1), owner i) and 12 edge records at
 * edges[12i..12i+11] (28 bytes each, svo_edge: both endpoints, absolute,
 * in the reference's edge order, owner i).  Host buffers. */
int svo_k_ref_vertex_edge(int device, const float* origin, const float* width, uint64_t n_nodes,
                          void* vertices, void* edges);

/* Voxelizes one mesh: writes up to cap hit codes (uint32 for key_bits 32,
 * uint64 for 64) in emission order, duplicates included; n_hits receives the
 * number of hits (may exceed cap). */
int svo_k_voxelize(int device, const float* xyz, uint64_t n_points, const int32_t* tri,
                   uint64_t n_tris, const int32_t* res, uint32_t key_bits, void* codes,
                   uint64_t cap, uint64_t* n_hits);

#ifdef __cplusplus
}
#endif
#endif /* SVO_SVO_KERNELS_H */
