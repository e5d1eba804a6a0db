// svo_main.cpp -- command line with the reference's interface (main.cpp:7-17):
//   svo [mesh.obj=bunny.obj] [res=64] [--intended] [--key-bits 32|64] [--out-dir DIR]
//       [--write-voxels]
// Builds the octree on the GPU and writes <stem>_<depth>_tree.obj.  With
// --write-voxels it also writes <stem>_1_voxel.obj, as the reference's NDEBUG
// build does from inside create() (SVO.cu:411-419), and prints its message
// first, in the same order.
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>

#include "SparseVoxelOctree.h"

int main(int argc, char** argv) {
  std::string mesh = "bunny.obj";
  int res = 64;
  uint32_t sem = SVO_SEM_REFERENCE, key_bits = 0;
  std::string out_dir;
  bool voxels = false;
  int pos = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--intended")) sem = SVO_SEM_INTENDED;
    else if (!std::strcmp(argv[i], "--key-bits") && i + 1 < argc) key_bits = std::stoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--out-dir") && i + 1 < argc) out_dir = argv[++i];
    else if (!std::strcmp(argv[i], "--write-voxels")) voxels = true;
    else if (pos == 0) { mesh = argv[i]; ++pos; }
    else if (pos == 1) { res = std::stoi(argv[i]); ++pos; }
  }
  try {
    svo_amd::SparseVoxelOctree svo(mesh, res, res, res, sem, key_bits);
    svo.create();
    const std::string stem = std::filesystem::path(mesh).stem().string();
    if (voxels) {
      const std::string vpath = svo.writeVoxel(stem, out_dir);
      std::fprintf(stdout, "[I/O] Writing data in obj voxels format to file %s \n",
                   vpath.c_str());  // SVO.cu:762
    }
    std::printf("Sparse Voxel Octree depth: %d\n", svo.depth());  // SVO.cu:488
    std::string path = out_dir.empty() ? stem + "_" + std::to_string(svo.depth()) + "_tree.obj"
                                       : out_dir + "/" + stem + "_" +
                                             std::to_string(svo.depth()) + "_tree.obj";
    std::fprintf(stdout, "[I/O] Writing octree data in obj format to file %s \n", path.c_str());
    svo.writeTree(stem, out_dir);
  } catch (const svo_amd::SvoError& e) {
    std::fprintf(stderr, "svo: %s\n", e.what());
    return 1;
  }
  return 0;
}
