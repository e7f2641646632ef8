// C++ facade of libgmorpho_decimate.so under the reference's Decimator
// interface (Decimator/Decimator.h:36-78, Decimator.cpp:1240-1297): the
// proxy simplification that consumes a closing mesh. Header-only over the
// C-ABI of gmorpho_decimate.h; link with -lgmorpho_decimate.
//
// Behaviour kept from the reference:
//  * the short constructor uses the bounding (linear) constraints and no
//    features, with 3 HC smoothing iterations (Decimator.cpp:1246-1265 and
//    the VCGOptimSession default);
//  * use_features lifts the maximum-edge-length penalty (Decimator.cpp:506);
//    the per-vertex feature tags are carried through collapses there but
//    enter no cost, so they are accepted and not used;
//  * the scale grid is read as signed chars indexed x + res_x*(y + res_y*z),
//    voxel i at bbox_min + cell_size*i (bbox_max is not read);
//  * error() is the priority of the last valid collapse taken from the heap
//    (currMetric), num_faces() the live face count after Optimize();
//  * GetMesh sets the positions and triangles only;
//  * failures throw Decimator::Exception with the library's message.
#ifndef BROXY_AMD_DECIMATOR_HPP
#define BROXY_AMD_DECIMATOR_HPP

#include <string>
#include <vector>

#include "gmorpho.hpp"
#include "gmorpho_decimate.h"

class Decimator {
 public:
  class Exception {
   public:
    explicit Exception(const std::string& msg) : msg_(msg) {}
    const std::string& msg() const { return msg_; }

   private:
    std::string msg_;
  };

  Decimator() {}
  Decimator(const MorphoGraphics::Mesh& mesh, int target_num_faces, float target_error, float max_edge_length,
            const MorphoGraphics::Vec3f& bbox_min, const MorphoGraphics::Vec3f& bbox_max,
            const MorphoGraphics::Vec3<unsigned int>& res, float cell_size, char* scale_grid)
      : Decimator(mesh, target_num_faces, target_error, true, false, max_edge_length, bbox_min, bbox_max, res,
                  cell_size, scale_grid, std::vector<bool>(), 3) {}
  Decimator(const MorphoGraphics::Mesh& mesh, int target_num_faces, float target_error, bool use_linear_constraints,
            bool use_features, float max_edge_length, const MorphoGraphics::Vec3f& bbox_min,
            const MorphoGraphics::Vec3f& /*bbox_max*/, const MorphoGraphics::Vec3<unsigned int>& res,
            float cell_size, char* scale_grid, const std::vector<bool>& /*feature_taggs*/, int num_hc_iters = 3)
      : target_num_faces_(target_num_faces),
        target_error_(target_error),
        use_linear_constraints_(use_linear_constraints),
        use_features_(use_features) {
    const std::vector<MorphoGraphics::Vec3f>& P = mesh.P();
    const std::vector<MorphoGraphics::Mesh::Triangle>& T = mesh.T();
    std::vector<float> p(3 * P.size());
    std::vector<uint32_t> t(3 * T.size());
    for (size_t i = 0; i < P.size(); ++i)
      for (int a = 0; a < 3; ++a) p[3 * i + a] = P[i][a];
    for (size_t i = 0; i < T.size(); ++i)
      for (int a = 0; a < 3; ++a) t[3 * i + a] = T[i][a];
    const float bmin[3] = {bbox_min[0], bbox_min[1], bbox_min[2]};
    const int r3[3] = {(int)res[0], (int)res[1], (int)res[2]};
    const int8_t* grid = use_features ? nullptr : reinterpret_cast<const int8_t*>(scale_grid);
    const int rc = gmd_create(p.data(), (int)P.size(), t.data(), (int)T.size(), target_num_faces, target_error,
                              use_linear_constraints ? 1 : 0, max_edge_length, bmin, r3, cell_size, grid,
                              num_hc_iters, &s_);
    if (rc != 0) {
      const std::string msg = s_ ? gmd_last_error(s_) : "Decimator Error: the session could not be created";
      if (s_) gmd_destroy(s_);
      s_ = nullptr;
      throw Exception(msg);
    }
    refresh();
  }
  Decimator(const Decimator&) = delete;
  Decimator& operator=(const Decimator&) = delete;
  virtual ~Decimator() {
    if (s_) gmd_destroy(s_);
  }

  int num_faces() { return num_faces_; }
  int target_num_faces() { return target_num_faces_; }
  float error() { return error_; }
  float target_error() { return target_error_; }
  bool use_linear_constraints() { return use_linear_constraints_; }
  bool use_features() { return use_features_; }
  void set_num_faces(int num_faces) { num_faces_ = num_faces; }
  void set_target_num_faces(int target_num_faces) { target_num_faces_ = target_num_faces; }
  void set_target_error(float target_error) { target_error_ = target_error; }
  void set_error(float error) { error_ = error; }
  void set_use_linear_constraints(float v) { use_linear_constraints_ = v != 0.f; }
  void set_use_features(float v) { use_features_ = v != 0.f; }

  // Decimator::Optimize (Decimator.cpp:1292-1297): collapses until a stop
  // rule holds; true while collapses remain queued.
  bool Optimize() {
    if (!s_) throw Exception("Decimator Error: no session");
    int more = 0;
    if (gmd_optimize(s_, 0.0, &more) != 0) throw Exception(gmd_last_error(s_));
    refresh();
    return more != 0;
  }
  // Decimator::GetMesh (Decimator.cpp:1203-1239): live vertices and faces,
  // re-indexed; like the reference it sets P and T only (N is left as it
  // was: call recomputeNormals() for normals of the proxy).
  void GetMesh(MorphoGraphics::Mesh& out) {
    if (!s_) throw Exception("Decimator Error: no session");
    int nf = 0, nv = 0;
    float err = 0.f;
    if (gmd_state(s_, &nf, &nv, &err) != 0) throw Exception(gmd_last_error(s_));
    std::vector<float> p(3 * (size_t)nv);
    std::vector<uint32_t> t(3 * (size_t)nf);
    if (gmd_get_mesh(s_, p.data(), t.data()) != 0) throw Exception(gmd_last_error(s_));
    out.P().resize(nv);
    out.T().resize(nf);
    for (int i = 0; i < nv; ++i) out.P()[i] = MorphoGraphics::Vec3f(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    for (int i = 0; i < nf; ++i)
      out.T()[i] = MorphoGraphics::Mesh::Triangle(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
  }

 private:
  void refresh() {
    int nf = 0, nv = 0;
    float err = 0.f;
    if (gmd_state(s_, &nf, &nv, &err) != 0) throw Exception(gmd_last_error(s_));
    num_faces_ = nf;
    error_ = err;
  }

  gmd_session* s_ = nullptr;
  int num_faces_ = 0, target_num_faces_ = 0;
  float error_ = 0.f, target_error_ = 0.f;
  bool use_linear_constraints_ = true, use_features_ = false;
};

#endif  // BROXY_AMD_DECIMATOR_HPP
