// obj_io.cpp -- OBJ mesh input and <stem>_<depth>_tree.obj output.
//
// Input follows BaseModel::loadOBJ (BaseModel.cpp:3-25), which reads with
// `while (!in.getline(buffer, 255).eof())`:
//   * a line is processed only if it ends with '\n' (the last line of a file
//     without a trailing newline hits end-of-file and is dropped);
//   * a line of 255 or more characters sets failbit; the reference then loops
//     forever, here it is SVO_E_PARSE;
//   * "v x y z" lines give points; "f a//b c//d e//f", then "f a/b c/d e/f",
//     then "f a b c" give the first three indices minus one; anything else,
//     including "f a/b/c ..." faces, is ignored.
// The file is read in one piece and its lines are parsed in parallel chunks
// that are concatenated in file order.
//
// Output follows SparseVoxelOctree::writeTree (SVO.cu:737-753) and write_cube
// (IO.h:23-90): per node eight "v x y z" lines (std::ostream float output is
// printf "%g" of the value promoted to double) and twelve "l a b" lines,
// indices offset by 8 per node.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "svo_host.h"

namespace svo {

namespace {

struct ParsedChunk {
  std::vector<float> pts;
  std::vector<int32_t> tris;
  int status = SVO_OK;
};

void parse_lines(const char* beg, const char* end, ParsedChunk& out) {
  char buf[256];
  const char* p = beg;
  while (p < end) {
    const char* nl = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    const size_t len = (size_t)((nl ? nl : end) - p);
    if (len >= 255) {
      out.status = SVO_E_PARSE;
      return;
    }
    if (!nl) return;  // no newline: end-of-file line, dropped
    std::memcpy(buf, p, len);
    buf[len] = 0;
    p = nl + 1;
    if (buf[0] == 'v' && buf[1] == ' ') {
      float x, y, z;
      if (std::sscanf(buf, "v %f %f %f", &x, &y, &z) == 3) {
        out.pts.push_back(x);
        out.pts.push_back(y);
        out.pts.push_back(z);
      }
    } else if (buf[0] == 'f' && buf[1] == ' ') {
      int f0, g0, f1, g1, f2, g2;
      if (std::sscanf(buf, "f %d//%d %d//%d %d//%d", &f0, &g0, &f1, &g1, &f2, &g2) == 6 ||
          std::sscanf(buf, "f %d/%d %d/%d %d/%d", &f0, &g0, &f1, &g1, &f2, &g2) == 6 ||
          std::sscanf(buf, "f %d %d %d", &f0, &f1, &f2) == 3) {
        // index - 1 with the two's-complement wrap (the reference's int
        // subtraction overflows for INT_MIN; its compiled code wraps)
        out.tris.push_back((int32_t)((uint32_t)f0 - 1u));
        out.tris.push_back((int32_t)((uint32_t)f1 - 1u));
        out.tris.push_back((int32_t)((uint32_t)f2 - 1u));
      }
    }
  }
}

unsigned n_workers(size_t bytes) {
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  hw = std::min(hw, 16u);
  const size_t per = 1 << 20;  // at least 1 MiB per worker
  return (unsigned)std::max<size_t>(1, std::min<size_t>(hw, bytes / per));
}

}  // namespace

int load_obj_file(const char* path, std::vector<float>& pts, std::vector<int32_t>& tris) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return SVO_E_IO;
  std::string data;
  {
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
      std::fclose(f);
      return SVO_E_IO;
    }
    data.resize((size_t)sz);
    if (sz && std::fread(&data[0], 1, (size_t)sz, f) != (size_t)sz) {
      std::fclose(f);
      return SVO_E_IO;
    }
    std::fclose(f);
  }
  const char* b = data.data();
  const char* e = b + data.size();
  const unsigned nw = n_workers(data.size());
  std::vector<const char*> cut(nw + 1);
  cut[0] = b;
  cut[nw] = e;
  for (unsigned i = 1; i < nw; ++i) {
    const char* c = b + data.size() * i / nw;
    if (c < cut[i - 1]) c = cut[i - 1];
    const char* nl = (const char*)std::memchr(c, '\n', (size_t)(e - c));
    cut[i] = nl ? nl + 1 : e;
  }
  std::vector<ParsedChunk> parts(nw);
  if (nw == 1) {
    parse_lines(cut[0], cut[1], parts[0]);
  } else {
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nw; ++i)
      th.emplace_back([&, i] { parse_lines(cut[i], cut[i + 1], parts[i]); });
    for (auto& t : th) t.join();
  }
  pts.clear();
  tris.clear();
  for (auto& p : parts) {
    if (p.status != SVO_OK) return p.status;
    pts.insert(pts.end(), p.pts.begin(), p.pts.end());
    tris.insert(tris.end(), p.tris.begin(), p.tris.end());
  }
  return SVO_OK;
}

namespace {

// Eight corners of write_cube (IO.h:133-140): origin + (0|w per axis), with
// the bare origin for v4; then twelve lines (IO.h:173-186).
void format_nodes(const float* origin, const float* width, uint64_t first, uint64_t end,
                  std::string& out) {
  static const int C[8][3] = {{0, 1, 1}, {0, 0, 1}, {0, 1, 0}, {0, 0, 0},
                              {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
  static const int Lp[12][2] = {{1, 2}, {2, 7}, {7, 8}, {8, 1}, {3, 4}, {4, 5},
                                {5, 6}, {6, 3}, {3, 1}, {4, 2}, {5, 7}, {6, 8}};
  char line[128];
  out.reserve(out.size() + (end - first) * 420);
  for (uint64_t t = first; t < end; ++t) {
    const float* o = origin + 3 * t;
    const float w = width[t];
    for (int k = 0; k < 8; ++k) {
      float v[3];
      for (int a = 0; a < 3; ++a) v[a] = (k == 3) ? o[a] : o[a] + (C[k][a] ? w : 0.0f);
      const int n = std::snprintf(line, sizeof line, "v %g %g %g\n", (double)v[0], (double)v[1],
                                  (double)v[2]);
      out.append(line, (size_t)n);
    }
    const uint64_t base = 8 * t;
    for (int k = 0; k < 12; ++k) {
      const int n = std::snprintf(line, sizeof line, "l %d %d\n", (int)(base + Lp[k][0]),
                                  (int)(base + Lp[k][1]));
      out.append(line, (size_t)n);
    }
  }
}

}  // namespace

int write_tree_obj(const char* path, const float* origin, const float* width, uint64_t n) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return SVO_E_IO;
  const uint64_t block = 1 << 16;  // nodes per formatting task
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  int st = SVO_OK;
  for (uint64_t b0 = 0; b0 < n && st == SVO_OK; b0 += block * hw) {
    const unsigned tasks = (unsigned)std::min<uint64_t>(hw, (n - b0 + block - 1) / block);
    std::vector<std::string> bufs(tasks);
    std::vector<std::thread> th;
    for (unsigned i = 0; i < tasks; ++i) {
      const uint64_t a = b0 + i * block, e = std::min(n, a + block);
      if (tasks == 1) format_nodes(origin, width, a, e, bufs[i]);
      else th.emplace_back([&, a, e, i] { format_nodes(origin, width, a, e, bufs[i]); });
    }
    for (auto& t : th) t.join();
    for (auto& s : bufs)
      if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) st = SVO_E_IO;
  }
  if (std::fclose(f) != 0) st = SVO_E_IO;
  return st;
}

}  // namespace svo
