// C++ facade of libgmorpho_hip.so under the reference's class and method
// names (namespace MorphoGraphics): a caller of GMorpho / ScaleField / Mesh
// swaps headers and links this library instead of the CUDA sources.
//
// Mirrors:
//   GMorpho            GMorpho/GMorpho.h:48-170 (Load, Update, Dilate*, Erode*,
//                      ExtractClosingMesh, bbox/res/cell_size, setters)
//   ScaleField         GMorpho/ScaleField.h:33-140 (Init, global scale,
//                      UpdateGridGlobalScale); per-voxel radii are given as a
//                      host int8 grid or an FP32 lattice instead of brush points
//   Mesh               Common/Mesh.h:47 (P, N, T vectors)
//   AxisAlignedBoundingBox  Common/BoundingVolume.h:104 (min, max)
// Errors surface as GMorpho::Exception ("GMorpho Error: ..."), as in the
// reference (GMorpho.h:52). Methods that the reference drives through OpenGL
// or that this build does not carry (frame-field dilation, bilateral normal
// filtering) throw instead of silently doing something else.
#ifndef GMORPHO_HPP
#define GMORPHO_HPP

#include <cmath>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "gmorpho_c.h"

namespace MorphoGraphics {

template <typename T>
class Vec3 {
 public:
  Vec3() : p_{T(0), T(0), T(0)} {}
  Vec3(T x, T y, T z) : p_{x, y, z} {}
  T& operator[](int i) { return p_[i]; }
  const T& operator[](int i) const { return p_[i]; }
  Vec3 operator*(T s) const { return Vec3(p_[0] * s, p_[1] * s, p_[2] * s); }
  friend Vec3 operator*(T s, const Vec3& v) { return v * s; }

 private:
  T p_[3];
};
typedef Vec3<float> Vec3f;
typedef Vec3<unsigned int> Vec3ui;
typedef Vec3<int> Vec3i;

class AxisAlignedBoundingBox {
 public:
  AxisAlignedBoundingBox() {}
  AxisAlignedBoundingBox(const Vec3f& mn, const Vec3f& mx) : min_(mn), max_(mx) {}
  const Vec3f& min() const { return min_; }
  const Vec3f& max() const { return max_; }

 private:
  Vec3f min_, max_;
};

class Mesh {
 public:
  typedef Vec3ui Triangle;
  std::vector<Vec3f>& P() { return P_; }
  std::vector<Vec3f>& N() { return N_; }
  std::vector<Triangle>& T() { return T_; }
  const std::vector<Vec3f>& P() const { return P_; }
  const std::vector<Vec3f>& N() const { return N_; }
  const std::vector<Triangle>& T() const { return T_; }
  void clear() {
    P_.clear();
    N_.clear();
    T_.clear();
  }

 private:
  std::vector<Vec3f> P_, N_;
  std::vector<Triangle> T_;
};

// Frame fields drive the rotated-cube dilation (DilateByRotCubeMipmap), which
// is not part of this build; the type exists so that Update's signature is
// the reference's.
class FrameField {};

class ScaleField {
 public:
  class Exception {
   public:
    explicit Exception(const std::string& msg) : msg_("ScaleField Error: " + msg) {}
    const std::string& msg() const { return msg_; }

   private:
    std::string msg_;
  };
  enum Mode { UNIFORM = 0, GRID_I8 = 1, LATTICE_F32 = 2 };

  float cell_size() const { return cell_size_; }
  const Vec3f& bbox_min() const { return bbox_min_; }
  const Vec3f& bbox_max() const { return bbox_max_; }
  const Vec3ui& res() const { return res_; }
  float global_scale() const { return global_scale_; }
  void set_global_scale(float global_scale) { global_scale_ = global_scale; }

  // ScaleField::Init (ScaleField.cu:61-71): res and cell_size of the fine grid
  // (2x the GMorpho grid, half its cell size, BroxyViewer.cpp:133-136).
  void Init(const Vec3f& bbox_min, const Vec3f& bbox_max, const Vec3ui& res, float cell_size) {
    bbox_min_ = bbox_min;
    bbox_max_ = bbox_max;
    res_ = res;
    cell_size_ = cell_size;
    mode_ = UNIFORM;
    grid_.clear();
    lattice_.clear();
  }
  // ScaleField::UpdateGridGlobalScale (ScaleField.cu:143-146): every fine
  // voxel gets floor(global_scale / cell_size).
  void UpdateGridGlobalScale() {
    mode_ = UNIFORM;
    grid_.clear();
    lattice_.clear();
  }
  // Per-fine-voxel radii (the reference's scale_grid_, ScaleField.h:135),
  // N^3 signed chars, x fastest.
  void SetScaleGrid(std::vector<int8_t> grid) {
    mode_ = GRID_I8;
    grid_ = std::move(grid);
  }
  // Radii sampled from an L^3 FP32 lattice (trilinear, floor, clamp [2,127]).
  void SetScaleLattice(std::vector<float> lattice, int L) {
    mode_ = LATTICE_F32;
    lattice_ = std::move(lattice);
    L_ = L;
  }
  Mode mode() const { return mode_; }
  const std::vector<int8_t>& scale_grid_host() const { return grid_; }
  const std::vector<float>& lattice() const { return lattice_; }
  int lattice_res() const { return L_; }
  int uniform_radius() const { return (int)std::floor(global_scale_ / cell_size_); }

 private:
  float global_scale_ = 0.f, cell_size_ = 1.f;
  Vec3f bbox_min_, bbox_max_;
  Vec3ui res_;
  Mode mode_ = UNIFORM;
  std::vector<int8_t> grid_;
  std::vector<float> lattice_;
  int L_ = 0;
};

class GMorpho {
 public:
  class Exception {
   public:
    explicit Exception(const std::string& msg) : msg_("GMorpho Error: " + msg) {}
    const std::string& msg() const { return msg_; }

   private:
    std::string msg_;
  };
  typedef AxisAlignedBoundingBox BoundingBox;

  explicit GMorpho(int device = 0) {
    if (gm_create(device, &ctx_) != GM_OK) {
      ctx_ = nullptr;
      throw Exception("no usable HIP device " + std::to_string(device));
    }
  }
  virtual ~GMorpho() {
    if (ctx_) gm_destroy(ctx_);
  }
  GMorpho(const GMorpho&) = delete;
  GMorpho& operator=(const GMorpho&) = delete;

  const BoundingBox& bbox() const { return bbox_; }
  const Vec3ui& res() const { return res_; }
  float cell_size() const { return cell_size_; }
  float se_size() const { return se_size_; }
  void set_se_size(float se_size) { se_size_ = se_size; }
  void set_use_asymmetric_closing(bool v) { use_asymmetric_closing_ = v; }
  bool use_asymmetric_closing() { return use_asymmetric_closing_; }
  void set_use_frame_field(bool v) { use_frame_field_ = v; }
  bool use_frame_field() { return use_frame_field_; }
  void set_bilateral_filtering(bool v) { bilateral_filtering_ = v; }
  Mesh get_morpho_mesh() { return morpho_mesh_; }
  gm_ctx* handle() { return ctx_; }

  // GMorpho::Load (GMorpho.cu:7731-7773).
  void Load(const float* P, const float* N, int num_of_vertices, const unsigned int* T, int num_of_faces,
            int base_res, float se_size) {
    se_size_ = se_size;
    check(gm_load(ctx_, P, N, num_of_vertices, T, num_of_faces, base_res, se_size));
    float mn[3], mx[3], cell;
    int res[3], dres[3];
    check(gm_grid_info(ctx_, mn, mx, res, dres, &cell));
    bbox_ = BoundingBox(Vec3f(mn[0], mn[1], mn[2]), Vec3f(mx[0], mx[1], mx[2]));
    res_ = Vec3ui(res[0], res[1], res[2]);
    cell_size_ = cell;
  }

  // GMorpho::Update (GMorpho.cu:7775-7817): dilation (cube when asymmetric
  // closing is on, else sphere) and erosion with the scale field, both
  // skipped when floor(2 * global_scale / cell_size) <= 1, then the mesh.
  void Update(Mesh& mesh, const ScaleField& scale_field, const FrameField& frame_field) {
    (void)frame_field;
    if (use_asymmetric_closing_ && use_frame_field_)
      throw Exception("frame-field (rotated cube) dilation is not part of this build");
    const int gate = (int)std::floor(2.f * scale_field.global_scale() / cell_size_);
    const int se = use_asymmetric_closing_ ? GM_SE_CUBE : GM_SE_SPHERE;
    if (gate > 1) {
      apply_scale(scale_field);
      check(gm_dilate(ctx_, se));
      check(gm_erode(ctx_));
    } else {
      check(gm_scale_uniform(ctx_, gate));
      check(gm_close(ctx_, se));  // radius <= 1: the closing is the input
    }
    ExtractClosingMesh(mesh, use_asymmetric_closing_ ? 200 : 20);
  }
  // GMorpho::Update (GMorpho.cu:7819-7845): cube dilation and sphere erosion
  // with the fixed radius floor(2 * se_size / cell_size) (Grid.cu:725).
  void Update(Mesh& mesh, float se_size) {
    se_size_ = se_size;
    check(gm_scale_uniform(ctx_, (int)std::floor(2.f * se_size / cell_size_)));
    check(gm_dilate(ctx_, GM_SE_CUBE));
    check(gm_erode(ctx_));
    ExtractClosingMesh(mesh);
  }
  void DilateBySphereMipmap(const ScaleField& scale_field) {
    apply_scale(scale_field);
    check(gm_dilate(ctx_, GM_SE_SPHERE));
  }
  void DilateByCubeMipmap(const ScaleField& scale_field) {
    apply_scale(scale_field);
    check(gm_dilate(ctx_, GM_SE_CUBE));
  }
  void ErodeBySphereMipmap(const ScaleField& scale_field) {
    apply_scale(scale_field);
    check(gm_erode(ctx_));
  }
  // GMorpho::ExtractClosingMesh (GMorpho.cu:6814-7001): marching cubes of the
  // closing grid with area-weighted unit normals, copied into mesh.
  void ExtractClosingMesh(Mesh& mesh, int num_bilateral_iters = 20) {
    (void)num_bilateral_iters;
    if (bilateral_filtering_) throw Exception("bilateral normal filtering is not part of this build");
    check(gm_extract_mesh(ctx_));
    uint32_t nv = 0, nt = 0;
    check(gm_mesh_counts(ctx_, &nv, &nt));
    std::vector<float> V(3 * (size_t)nv), Nr(3 * (size_t)nv);
    std::vector<uint32_t> T(3 * (size_t)nt);
    check(gm_mesh_copy(ctx_, V.data(), Nr.data(), T.data()));
    mesh.clear();
    mesh.P().resize(nv);
    mesh.N().resize(nv);
    mesh.T().resize(nt);
    for (uint32_t i = 0; i < nv; ++i) {
      mesh.P()[i] = Vec3f(V[3 * i], V[3 * i + 1], V[3 * i + 2]);
      mesh.N()[i] = Vec3f(Nr[3 * i], Nr[3 * i + 1], Nr[3 * i + 2]);
    }
    for (uint32_t i = 0; i < nt; ++i) mesh.T()[i] = Vec3ui(T[3 * i], T[3 * i + 1], T[3 * i + 2]);
    morpho_mesh_ = mesh;
  }

 private:
  void check(int rc) const {
    if (rc != GM_OK) {
      std::string m = gm_last_error(ctx_);
      const std::string pre = "GMorpho Error: ";
      if (m.compare(0, pre.size(), pre) == 0) m = m.substr(pre.size());
      throw Exception(m);
    }
  }
  void apply_scale(const ScaleField& sf) {
    switch (sf.mode()) {
      case ScaleField::GRID_I8:
        check(gm_scale_grid_i8(ctx_, sf.scale_grid_host().data()));
        break;
      case ScaleField::LATTICE_F32:
        check(gm_scale_lattice_f32(ctx_, sf.lattice().data(), sf.lattice_res()));
        break;
      default:
        check(gm_scale_uniform(ctx_, sf.uniform_radius()));
    }
  }

  gm_ctx* ctx_ = nullptr;
  BoundingBox bbox_;
  Vec3ui res_;
  float cell_size_ = 0.f, se_size_ = 0.f;
  bool use_asymmetric_closing_ = false, use_frame_field_ = false, bilateral_filtering_ = false;
  Mesh morpho_mesh_;
};

}  // namespace MorphoGraphics

#endif
