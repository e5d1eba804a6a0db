// SparseVoxelOctree.cpp -- see SparseVoxelOctree.h.
#include "SparseVoxelOctree.h"

namespace svo_amd {

static void check(int st, const char* what) {
  if (st != SVO_OK) throw SvoError(st, what);
}

SparseVoxelOctree::SparseVoxelOctree(const std::string& filename, const int& grid_x,
                                     const int& grid_y, const int& grid_z, uint32_t semantics,
                                     uint32_t key_bits, int device) {
  const int g[3] = {grid_x, grid_y, grid_z};
  init(filename, g, semantics, key_bits, device);
}

SparseVoxelOctree::SparseVoxelOctree(const std::string& filename, const int grid[3],
                                     uint32_t semantics, uint32_t key_bits, int device) {
  init(filename, grid, semantics, key_bits, device);
}

void SparseVoxelOctree::init(const std::string& filename, const int grid[3], uint32_t semantics,
                             uint32_t key_bits, int device) {
  // BaseModel(filename): BaseModel.h:64-73
  check(svo_load_obj(filename.c_str(), &mesh_), ("loading obj " + filename).c_str());
  for (int a = 0; a < 3; ++a) params_.res[a] = grid[a];
  params_.semantics = semantics;
  params_.key_bits = key_bits;
  const int st = svo_ctx_create(device, 0, &ctx_);
  if (st != SVO_OK) {
    svo_mesh_free(mesh_);
    mesh_ = nullptr;
    throw SvoError(st, "creating the device context");
  }
}

SparseVoxelOctree::~SparseVoxelOctree() {
  if (tree_) svo_tree_free(tree_);
  if (mesh_) svo_mesh_free(mesh_);
  if (ctx_) svo_ctx_destroy(ctx_);
}

void SparseVoxelOctree::create() {
  if (tree_) {
    svo_tree_free(tree_);
    tree_ = nullptr;
  }
  svo_mesh_desc d;
  check(svo_mesh_view(mesh_, &d), "mesh view");
  check(svo_build(ctx_, &d, &params_, &tree_), "building the octree");
  check(svo_tree_info_get(tree_, &info_), "tree info");
}

std::string SparseVoxelOctree::writeTree(const std::string& base_filename,
                                         const std::string& dir) {
  if (!tree_) throw SvoError(SVO_E_INVALID, "writeTree before create");
  char path[4096];
  check(svo_write_obj(tree_, dir.c_str(), base_filename.c_str(), path, sizeof path),
        "writing the tree");
  return path;
}

std::string SparseVoxelOctree::writeVoxel(const std::string& base_filename,
                                          const std::string& dir) {
  if (!tree_) throw SvoError(SVO_E_INVALID, "writeVoxel before create");
  char path[4096];
  check(svo_write_voxels_obj(tree_, dir.c_str(), base_filename.c_str(), path, sizeof path),
        "writing the voxels");
  return path;
}

std::vector<svo_node_rec> SparseVoxelOctree::nodes() const {
  std::vector<svo_node_rec> v(info_.n_nodes);
  check(svo_tree_copy_field(tree_, SVO_FIELD_NODES_AOS, v.data(), v.size() * sizeof(svo_node_rec)),
        "copying nodes");
  return v;
}

std::vector<svo_vertex> SparseVoxelOctree::vertices() const {
  std::vector<svo_vertex> v(info_.n_vertices);
  check(svo_tree_copy_field(tree_, SVO_FIELD_VERTICES, v.data(), v.size() * sizeof(svo_vertex)),
        "copying vertices");
  return v;
}

std::vector<svo_edge> SparseVoxelOctree::edges() const {
  std::vector<svo_edge> v(info_.n_edges);
  check(svo_tree_copy_field(tree_, SVO_FIELD_EDGES, v.data(), v.size() * sizeof(svo_edge)),
        "copying edges");
  return v;
}

}  // namespace svo_amd
