// Drives the Decimator facade (include/decimator.hpp) the way the reference's
// application does after a closing (BroxyViewer.cpp: Decimator built on the
// closing mesh with the ScaleField grid, Optimize, GetMesh), and stores the
// proxy as OFF. Used by tests/test_cpu_decimate.py.
//
// usage: decimate_facade in.off out.off target_faces target_error max_edge_length
//                        short|full lin feat hc grid.raw|- rx ry rz bminx bminy bminz cell
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "decimator.hpp"

int main(int argc, char** argv) {
  if (argc != 18) {
    std::fprintf(stderr, "usage: %s in.off out.off target error maxlen short|full lin feat hc grid|- rx ry rz "
                 "bx by bz cell\n", argv[0]);
    return 2;
  }
  try {
    MorphoGraphics::Mesh mesh;
    mesh.load(argv[1]);
    const int target = std::atoi(argv[3]);
    const float err = (float)std::atof(argv[4]), maxlen = (float)std::atof(argv[5]);
    const bool shortCtor = std::string(argv[6]) == "short";
    const bool lin = std::atoi(argv[7]) != 0, feat = std::atoi(argv[8]) != 0;
    const int hc = std::atoi(argv[9]);
    const MorphoGraphics::Vec3<unsigned int> res((unsigned)std::atoi(argv[11]), (unsigned)std::atoi(argv[12]),
                                                 (unsigned)std::atoi(argv[13]));
    const MorphoGraphics::Vec3f bmin((float)std::atof(argv[14]), (float)std::atof(argv[15]),
                                     (float)std::atof(argv[16]));
    const float cell = (float)std::atof(argv[17]);
    std::vector<char> grid;
    char* g = nullptr;
    if (std::string(argv[10]) != "-") {
      std::ifstream in(argv[10], std::ios::binary);
      grid.assign((size_t)res[0] * res[1] * res[2], 0);
      in.read(grid.data(), (std::streamsize)grid.size());
      if (!in) throw MorphoGraphics::Mesh::Exception("short scale grid file");
      g = grid.data();
    }
    const MorphoGraphics::Vec3f bmax(bmin[0] + cell * res[0], bmin[1] + cell * res[1], bmin[2] + cell * res[2]);
    std::vector<bool> tags(mesh.P().size(), false);
    Decimator* d = shortCtor ? new Decimator(mesh, target, err, maxlen, bmin, bmax, res, cell, g)
                             : new Decimator(mesh, target, err, lin, feat, maxlen, bmin, bmax, res, cell, g, tags, hc);
    const int before = d->num_faces();
    const bool more = d->Optimize();
    MorphoGraphics::Mesh proxy;
    d->GetMesh(proxy);
    proxy.store(argv[2]);
    std::printf("faces %d -> %d error %.9g more %d linear %d features %d\n", before, d->num_faces(), d->error(),
                more ? 1 : 0, d->use_linear_constraints() ? 1 : 0, d->use_features() ? 1 : 0);
    delete d;
  } catch (const Decimator::Exception& e) {
    std::fprintf(stderr, "%s\n", e.msg().c_str());
    return 1;
  } catch (const MorphoGraphics::Mesh::Exception& e) {
    std::fprintf(stderr, "%s\n", e.msg().c_str());
    return 1;
  }
  return 0;
}
