// Drives the C++ facade (include/gmorpho.hpp) the way the reference's viewer
// drives GMorpho (BroxyViewer.cpp:126-157): Load, ScaleField Init + global
// scale, Update -> Mesh. Writes the closing grid (cell bytes) and the
// triangle list to files so that a test can compare them with the oracle's
// record. Usage: facade_closing <out_prefix> [seed]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gmorpho.hpp"
#include "gmorpho_meshgen.h"

using namespace MorphoGraphics;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <out_prefix> [seed]\n", argv[0]);
    return 2;
  }
  const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 2017ull;
  int nv = 0, nf = 0;
  if (gmg_icosphere_counts(3, &nv, &nf)) return 1;
  const int n = 64;
  std::vector<float> P(3 * (size_t)nv * n);
  std::vector<uint32_t> T(3 * (size_t)nf * n);
  if (gmg_sphere_soup(n, 3, seed, 0.15, 0.85, 0.03, 0.12, P.data(), T.data())) return 1;
  try {
    GMorpho gmorpho(0);
    gmorpho.Load(P.data(), nullptr, nv * n, T.data(), nf * n, 64, 0.3f);
    ScaleField sf;
    Vec3ui res = 2u * gmorpho.res();
    float cell = 0.5f * gmorpho.cell_size();
    // radius 4 fine voxels: global scale 4 fine cells (plus a little, floor)
    sf.set_global_scale(4.25f * cell);
    sf.Init(gmorpho.bbox().min(), gmorpho.bbox().max(), res, cell);
    sf.UpdateGridGlobalScale();
    gmorpho.set_use_asymmetric_closing(false);
    Mesh mesh;
    gmorpho.Update(mesh, sf, FrameField());
    const int R = (int)gmorpho.res()[0];
    std::vector<uint8_t> C((size_t)R * R * R);
    if (gm_grid_get_cells(gmorpho.handle(), GM_GRID_CLOSING, C.data())) return 1;
    std::string pre = argv[1];
    FILE* f = std::fopen((pre + "_closing.u8").c_str(), "wb");
    std::fwrite(C.data(), 1, C.size(), f);
    std::fclose(f);
    f = std::fopen((pre + "_tris.u32").c_str(), "wb");
    for (auto& t : mesh.T())
      for (int k = 0; k < 3; ++k) {
        uint32_t v = t[k];
        std::fwrite(&v, 4, 1, f);
      }
    std::fclose(f);
    std::printf("%zu %zu\n", mesh.P().size(), mesh.T().size());
  } catch (const GMorpho::Exception& e) {
    std::fprintf(stderr, "%s\n", e.msg().c_str());
    return 1;
  }
  return 0;
}
