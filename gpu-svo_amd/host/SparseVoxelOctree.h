// SparseVoxelOctree.h -- C++ host interface with the shape of the reference's
// class (SVO.h:39-89): construct from an OBJ file and a grid size, create(),
// writeTree(stem).  It runs on the MI355X library through the C ABI
// (include/svo/svo.h) and adds read-outs the reference keeps private.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "svo/svo.h"

namespace svo_amd {

class SvoError : public std::runtime_error {
 public:
  SvoError(int status, const std::string& what)
      : std::runtime_error(what + ": " + svo_status_str(status)), status(status) {}
  int status;
};

class SparseVoxelOctree {
 public:
  // SVO.h:61-67
  SparseVoxelOctree(const std::string& filename, const int& grid_x, const int& grid_y,
                    const int& grid_z, uint32_t semantics = SVO_SEM_REFERENCE,
                    uint32_t key_bits = 0, int device = 0);
  // SVO.h:56-59 (Eigen::Vector3i replaced by three ints)
  SparseVoxelOctree(const std::string& filename, const int grid[3],
                    uint32_t semantics = SVO_SEM_REFERENCE, uint32_t key_bits = 0,
                    int device = 0);
  ~SparseVoxelOctree();
  SparseVoxelOctree(const SparseVoxelOctree&) = delete;
  SparseVoxelOctree& operator=(const SparseVoxelOctree&) = delete;

  // SVO.cu:341-520 (+ constructNodeAtrributes).  Throws SvoError.
  void create();
  // SVO.cu:737-753: writes <base_filename>_<depth>_tree.obj; returns the path.
  std::string writeTree(const std::string& base_filename, const std::string& dir = "");
  // writeVoxel (SVO.cu:755-777): <base_filename>_1_voxel.obj
  std::string writeVoxel(const std::string& base_filename, const std::string& dir = "");

  int depth() const { return info_.depth; }
  const svo_tree_info& info() const { return info_; }
  const svo_tree* tree() const { return tree_; }
  std::vector<svo_node_rec> nodes() const;
  std::vector<svo_vertex> vertices() const;
  std::vector<svo_edge> edges() const;

 private:
  void init(const std::string& filename, const int grid[3], uint32_t semantics,
            uint32_t key_bits, int device);
  svo_ctx* ctx_ = nullptr;
  svo_mesh* mesh_ = nullptr;
  svo_tree* tree_ = nullptr;
  svo_params params_{};
  svo_tree_info info_{};
};

}  // namespace svo_amd
