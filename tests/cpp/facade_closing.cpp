// Drives the C++ facade (include/gmorpho.hpp) the way the reference's viewer
// drives GMorpho (BroxyViewer.cpp:126-157): Load, ScaleField Init + global
// scale, Update -> Mesh. Writes the closing grid (cell bytes), the triangle
// list, the mesh as OFF and as PN (positions + normals) so that a test can
// compare them with the oracle's, then reads the OFF file back.
// Usage: facade_closing <out_prefix> [seed] [mode]
//   mode sphere (default): ball closing, no normal filtering;
//   viewer: the viewer's defaults, asymmetric (cube) closing with bilateral
//           normal filtering (BroxyViewer.cpp:66, 125, 138), 200 iterations;
//   fixed:  Update(Mesh&, float se_size) (GMorpho.cu:7819-7845);
//   brush:  ball closing with a field edited by brush strokes (AddPoint +
//           UpdateGrid, BroxyViewer.cpp:670-671), the radius grid written too
//           (ScaleField::GetScaleGridCPU), then the viewer's proxy: the
//           Decimator loop on the closing mesh with that grid
//           (BroxyViewer.cpp:1060-1125), stored as <out_prefix>_proxy.off.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "decimator.hpp"
#include "gmorpho.hpp"
#include "gmorpho_meshgen.h"

using namespace MorphoGraphics;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <out_prefix> [seed]\n", argv[0]);
    return 2;
  }
  const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 2017ull;
  const std::string mode = argc > 3 ? argv[3] : "sphere";
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
    Mesh mesh;
    if (mode == "viewer") {
      gmorpho.set_use_asymmetric_closing(true);
      gmorpho.set_bilateral_filtering(true);
      gmorpho.Update(mesh, sf, FrameField());
    } else if (mode == "brush") {
      gmorpho.set_use_asymmetric_closing(false);
      sf.AddPoint(Vec3f(0.5f, 0.5f, 0.5f), 0.12f, 0.2f);
      sf.UpdateGrid();
      gmorpho.Update(mesh, sf, FrameField());
      sf.AddPoint(Vec3f(0.35f, 0.6f, 0.45f), 0.05f, 0.09f);
      sf.points().back().set_edit_mode(MIN_BRUSH);
      sf.UpdateGrid();
      gmorpho.Update(mesh, sf, FrameField());
      const size_t n = (size_t)res[0] * res[1] * res[2];
      char* g = nullptr;
      sf.GetScaleGridCPU(&g);
      FILE* fg = std::fopen((std::string(argv[1]) + "_scale.i8").c_str(), "wb");
      std::fwrite(g, 1, n, fg);
      std::fclose(fg);
      // the viewer's proxy (ball closing: target error 0.001, 20 faces,
      // edge-length alpha 1, linear constraints, no features, 3 HC iterations)
      std::vector<bool> tags(mesh.P().size(), false);
      Decimator decimator(mesh, 20, 0.001f, true, false, 1.f, sf.bbox_min(), sf.bbox_max(), sf.res(),
                          sf.cell_size(), g, tags, 3);
      bool more = true;
      while (more && decimator.num_faces() > decimator.target_num_faces() &&
             decimator.error() < decimator.target_error())
        more = decimator.Optimize();
      Mesh proxy;
      decimator.GetMesh(proxy);
      proxy.store(std::string(argv[1]) + "_proxy.off");
      delete[] g;
    } else if (mode == "fixed") {
      // se_size in world units: radius floor(2 * se_size / cell_size)
      gmorpho.Update(mesh, 4.25f * cell);
    } else {
      gmorpho.set_use_asymmetric_closing(false);
      gmorpho.Update(mesh, sf, FrameField());
    }
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
    mesh.store(pre + "_mesh.off");
    mesh.store(pre + "_mesh.pn");
    Mesh back;
    back.load(pre + "_mesh.off");
    bool same = back.P().size() == mesh.P().size() && back.T().size() == mesh.T().size();
    for (size_t i = 0; same && i < mesh.T().size(); ++i)
      for (int k = 0; k < 3; ++k) same = same && back.T()[i][k] == mesh.T()[i][k];
    if (!same) {
      std::fprintf(stderr, "OFF round trip differs\n");
      return 1;
    }
    std::printf("%zu %zu\n", mesh.P().size(), mesh.T().size());
  } catch (const GMorpho::Exception& e) {
    std::fprintf(stderr, "%s\n", e.msg().c_str());
    return 1;
  } catch (const Mesh::Exception& e) {
    std::fprintf(stderr, "%s\n", e.msg().c_str());
    return 1;
  } catch (const ScaleField::Exception& e) {
    std::fprintf(stderr, "%s\n", e.msg().c_str());
    return 1;
  } catch (const Decimator::Exception& e) {
    std::fprintf(stderr, "%s\n", e.msg().c_str());
    return 1;
  }
  return 0;
}
