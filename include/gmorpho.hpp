// C++ facade of libgmorpho_hip.so under the reference's class and method
// names (namespace MorphoGraphics): a caller of GMorpho / ScaleField / Mesh
// swaps headers and links this library instead of the CUDA sources.
//
// Mirrors:
//   GMorpho            GMorpho/GMorpho.h:48-170 (Load, Update, Dilate*, Erode*,
//                      ExtractClosingMesh, bbox/res/cell_size, setters)
//   ScaleField         GMorpho/ScaleField.h:33-140 (Init, global scale,
//                      UpdateGridGlobalScale, AddPoint + UpdateGrid brush
//                      strokes applied on the device); per-voxel radii may
//                      also be given as a host int8 grid or an FP32 lattice
//   Mesh               Common/Mesh.h:47 (P, N, T vectors; OFF and PN load and
//                      store, 580-662; recomputeNormals, 297-306)
//   AxisAlignedBoundingBox  Common/BoundingVolume.h:104 (min, max)
// Errors surface as GMorpho::Exception ("GMorpho Error: ..."), as in the
// reference (GMorpho.h:52). The frame-field (rotated cube) dilation, which
// this build does not carry, throws instead of silently doing something else.
#ifndef GMORPHO_HPP
#define GMORPHO_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
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
  class Exception {
   public:
    explicit Exception(const std::string& msg) : msg_(msg) {}
    const std::string& msg() const { return msg_; }

   private:
    std::string msg_;
  };
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

  // Mesh::recomputeNormals (Common/Mesh.h:297-306): per vertex, the sum of
  // the unit normals of its triangles (in triangle order), normalized; a zero
  // vector stays zero.
  void recomputeNormals() {
    N_.assign(P_.size(), Vec3f());
    for (const Triangle& t : T_) {
      if (t[0] >= P_.size() || t[1] >= P_.size() || t[2] >= P_.size()) continue;
      Vec3f n = unit(cross(sub(P_[t[1]], P_[t[0]]), sub(P_[t[2]], P_[t[0]])));
      for (int j = 0; j < 3; ++j)
        for (int a = 0; a < 3; ++a) N_[t[j]][a] += n[a];
    }
    for (Vec3f& n : N_) n = unit(n);
  }

  // Mesh::load / store (Common/Mesh.h:580-600): by the file name's suffix.
  void load(const std::string& filename) {
    const std::string f = suffix(filename);
    if (f == "off") loadOFF(filename);
    else if (f == ".pn") loadPN(filename);
    else throw Exception("Error loading mesh: unrecognized file format.");
  }
  void store(const std::string& filename) const {
    const std::string f = suffix(filename);
    if (f == "off") storeOFF(filename);
    else if (f == ".pn") storePN(filename);
    else throw Exception("Error storing mesh: unrecognized file format.");
  }

  // Mesh::loadOFF (Common/Mesh.h:602-622): "OFF", counts, positions, then
  // faces of s vertices each, fan-triangulated as (v0, v[j-1], v[j]); then
  // recomputeNormals. Indices out of range and truncated files are errors.
  void loadOFF(const std::string& filename) {
    clear();
    std::ifstream in(filename.c_str());
    if (!in) throw Exception("Error loading OFF file: " + filename);
    std::string head;
    unsigned long long nv = 0, nf = 0, ne = 0;
    in >> head >> nv >> nf >> ne;
    if (!in || head != "OFF") throw Exception("Error loading OFF file: bad header in " + filename);
    P_.resize(nv);
    for (auto& p : P_) in >> p[0] >> p[1] >> p[2];
    for (unsigned long long i = 0; i < nf && in; ++i) {
      unsigned int sz = 0;
      in >> sz;
      std::vector<unsigned int> v(sz);
      for (auto& x : v) in >> x;
      for (unsigned int x : v)
        if (x >= nv) throw Exception("Error loading OFF file: vertex index out of range in " + filename);
      for (unsigned int j = 2; j < sz; ++j) T_.push_back(Triangle(v[0], v[j - 1], v[j]));
    }
    if (!in) throw Exception("Error loading OFF file: truncated " + filename);
    recomputeNormals();
  }
  // Mesh::storeOFF (Common/Mesh.h:624-636).
  void storeOFF(const std::string& filename) const {
    std::ofstream out(filename.c_str());
    if (!out) throw Exception("Error storing OFF file: " + filename);
    out.precision(9);
    out << "OFF\n" << P_.size() << " " << T_.size() << " 0\n";
    for (const auto& p : P_) out << p[0] << " " << p[1] << " " << p[2] << "\n";
    for (const auto& t : T_) out << "3 " << t[0] << " " << t[1] << " " << t[2] << "\n";
    if (!out) throw Exception("Error storing OFF file: " + filename);
  }
  // Mesh::loadPN / storePN (Common/Mesh.h:638-672): raw float32 records
  // (px py pz nx ny nz) per point, no triangles.
  void loadPN(const std::string& filename) {
    clear();
    FILE* f = std::fopen(filename.c_str(), "rb");
    if (!f) throw Exception("Error loading a PN file: Cannot read file" + filename);
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    const size_t n = bytes > 0 ? (size_t)bytes / (6 * sizeof(float)) : 0;
    std::vector<float> pn(6 * n);
    const size_t got = n ? std::fread(pn.data(), sizeof(float), 6 * n, f) : 0;
    std::fclose(f);
    if (got != 6 * n) throw Exception("Error loading a PN file: Cannot read file" + filename);
    P_.resize(n);
    N_.resize(n);
    for (size_t i = 0; i < n; ++i) {
      P_[i] = Vec3f(pn[6 * i], pn[6 * i + 1], pn[6 * i + 2]);
      N_[i] = Vec3f(pn[6 * i + 3], pn[6 * i + 4], pn[6 * i + 5]);
    }
  }
  void storePN(const std::string& filename) const {
    FILE* f = std::fopen(filename.c_str(), "wb");
    if (!f) throw Exception("Error storing a PN file: Cannot open file" + filename);
    for (size_t i = 0; i < P_.size(); ++i) {
      const Vec3f n = i < N_.size() ? N_[i] : Vec3f();
      const float buf[6] = {P_[i][0], P_[i][1], P_[i][2], n[0], n[1], n[2]};
      std::fwrite(buf, sizeof(float), 6, f);
    }
    std::fclose(f);
  }

 private:
  static std::string suffix(const std::string& f) { return f.size() >= 3 ? f.substr(f.size() - 3) : f; }
  static Vec3f sub(const Vec3f& a, const Vec3f& b) { return Vec3f(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
  static Vec3f cross(const Vec3f& a, const Vec3f& b) {
    return Vec3f(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
  }
  // Vec3::normalize (Common/Vec3.h:208-217): times 1/length, zero stays zero.
  static Vec3f unit(const Vec3f& a) {
    const float l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (l == 0.f) return a;
    const float inv = 1.f / l;
    return Vec3f(a[0] * inv, a[1] * inv, a[2] * inv);
  }
  std::vector<Vec3f> P_, N_;
  std::vector<Triangle> T_;
};

// Frame fields drive the rotated-cube dilation (DilateByRotCubeMipmap), which
// is not part of this build; the type exists so that Update's signature is
// the reference's.
class FrameField {};

// EditMode (ScaleField.h:32).
typedef enum { NO_EDITION = GM_BRUSH_NONE, MAX_BRUSH = GM_BRUSH_MAX, MIN_BRUSH = GM_BRUSH_MIN,
               ZERO_BRUSH = GM_BRUSH_ZERO } EditMode;

class ScaleField {
 public:
  class Exception {
   public:
    explicit Exception(const std::string& msg) : msg_("ScaleField Error: " + msg) {}
    const std::string& msg() const { return msg_; }

   private:
    std::string msg_;
  };
  // ScaleField::ScalePoint (ScaleField.h:43-80).
  class ScalePoint {
   public:
    ScalePoint() {}
    ScalePoint(const Vec3f& position, float scale) : position_(position), scale_(scale), support_(scale) {}
    ScalePoint(const Vec3f& position, float scale, float support)
        : position_(position), scale_(scale), support_(support) {}
    const Vec3f& position() const { return position_; }
    Vec3f& position() { return position_; }
    float scale() const { return scale_; }
    float support() const { return support_; }
    EditMode edit_mode() const { return edit_mode_; }
    void set_position(const Vec3f& position) { position_ = position; }
    void set_scale(float scale) { scale_ = scale; }
    void set_support(float support) { support_ = support; }
    void set_edit_mode(EditMode edit_mode) { edit_mode_ = edit_mode; }

   private:
    Vec3f position_;
    float scale_ = 0.f, support_ = 0.f;
    EditMode edit_mode_ = MAX_BRUSH;
  };
  enum Mode { UNIFORM = 0, GRID_I8 = 1, LATTICE_F32 = 2, BRUSHED = 3 };

  ScaleField() : id_(next_id()) {}
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
    reset(UNIFORM);
  }
  // ScaleField::UpdateGridGlobalScale (ScaleField.cu:143-146): every fine
  // voxel gets floor(global_scale / cell_size).
  void UpdateGridGlobalScale() { reset(UNIFORM); }
  // ScaleField::AddPoint (ScaleField.h:102-109).
  void AddPoint(const Vec3f& position, float scale) { points_.push_back(ScalePoint(position, scale)); }
  void AddPoint(const Vec3f& position, float scale, float support) {
    points_.push_back(ScalePoint(position, scale, support));
  }
  const std::vector<ScalePoint>& points() const { return points_; }
  std::vector<ScalePoint>& points() { return points_; }
  int GetNumberOfPoints() const { return (int)points_.size(); }
  // ScaleField::UpdateGrid (ScaleField.cu:148-192): the last point is
  // brushed into the grid (on the device, at the next GMorpho call that
  // reads the field; strokes accumulate until UpdateGridGlobalScale).
  void UpdateGrid() {
    if (points_.empty()) return;
    if (mode_ != BRUSHED) reset(BRUSHED);
    strokes_.push_back(points_.back());
  }
  // Per-fine-voxel radii (the reference's scale_grid_, ScaleField.h:135),
  // N^3 signed chars, x fastest.
  void SetScaleGrid(std::vector<int8_t> grid) {
    reset(GRID_I8);
    grid_ = std::move(grid);
  }
  // Radii sampled from an L^3 FP32 lattice (trilinear, floor, clamp [2,127]).
  void SetScaleLattice(std::vector<float> lattice, int L) {
    reset(LATTICE_F32);
    lattice_ = std::move(lattice);
    L_ = L;
  }
  Mode mode() const { return mode_; }
  const std::vector<int8_t>& scale_grid_host() const { return grid_; }
  const std::vector<float>& lattice() const { return lattice_; }
  int lattice_res() const { return L_; }
  int uniform_radius() const { return (int)std::floor(global_scale_ / cell_size_); }
  // Brush strokes since the last reset, and which field state they belong to.
  const std::vector<ScalePoint>& strokes() const { return strokes_; }
  unsigned long long id() const { return id_; }
  unsigned long long generation() const { return generation_; }
  // ScaleField::GetScaleGridCPU (ScaleField.cu:74-79): a new[] copy of the
  // res^3 grid, which the caller deletes (the viewer hands it to the
  // Decimator). A brushed field lives on the device of the GMorpho that last
  // applied it (GMorpho::Update); a uniform field is the memset byte of
  // UpdateGridGlobalScale; a lattice field has no int8 grid here.
  void GetScaleGridCPU(char** scale_grid_cpu_ptr) const {
    const size_t n = (size_t)res_[0] * res_[1] * res_[2];
    if (mode_ == LATTICE_F32) throw Exception("a lattice field has no int8 scale grid");
    if (mode_ == BRUSHED && !device_) throw Exception("brush strokes reach the grid at GMorpho::Update first");
    if (mode_ == GRID_I8 && grid_.size() != n) throw Exception("the scale grid does not match res");
    char* g = new char[n];
    if (mode_ == UNIFORM) {
      std::memset(g, uniform_radius() & 0xff, n);
    } else if (mode_ == GRID_I8) {
      std::memcpy(g, grid_.data(), n);
    } else if (gm_scale_grid_get(device_, reinterpret_cast<int8_t*>(g)) != GM_OK) {
      const std::string m = gm_last_error(device_);
      delete[] g;
      throw Exception(m);
    }
    *scale_grid_cpu_ptr = g;
  }
  // the context that last applied this field (set by GMorpho)
  void set_device(gm_ctx* ctx) const { device_ = ctx; }

 private:
  static unsigned long long next_id() {
    static unsigned long long n = 0;
    return ++n;
  }
  void reset(Mode m) {
    mode_ = m;
    grid_.clear();
    lattice_.clear();
    strokes_.clear();
    ++generation_;
  }
  float global_scale_ = 0.f, cell_size_ = 1.f;
  Vec3f bbox_min_, bbox_max_;
  Vec3ui res_;
  Mode mode_ = UNIFORM;
  std::vector<int8_t> grid_;
  std::vector<float> lattice_;
  int L_ = 0;
  std::vector<ScalePoint> points_, strokes_;
  unsigned long long id_ = 0, generation_ = 0;
  mutable gm_ctx* device_ = nullptr;
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

  // GMorpho::Load with the voxelizer's MultiSlicing in place of the
  // conservative voxelization (the alternative left commented out at
  // GMorpho.cu:7741; Voxelizer::VoxelizeByMultiSlicing, Voxelizer.cu:375-400),
  // with GMorpho::Load's margin.
  void LoadByMultiSlicing(const float* P, const float* N, int num_of_vertices, const unsigned int* T,
                          int num_of_faces, int base_res, float se_size) {
    se_size_ = se_size;
    check(gm_load_mesh(ctx_, P, N, num_of_vertices, T, num_of_faces));
    check(gm_voxelize_multislice(ctx_, base_res, se_size + 2.f / (2 * (float)base_res)));
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
    if (gate > 1)
      apply_scale(scale_field);
    else
      check(gm_scale_uniform(ctx_, gate));  // radius <= 1: the closing is the input
    // dilation, erosion and marching cubes enqueued as one sequence
    check(gm_update(ctx_, se));
    ReadClosingMesh(mesh, use_asymmetric_closing_ ? 200 : 20);
  }
  // GMorpho::Update (GMorpho.cu:7819-7845): cube dilation and sphere erosion
  // with the fixed radius floor(2 * se_size / cell_size) (Grid.cu:725).
  void Update(Mesh& mesh, float se_size) {
    se_size_ = se_size;
    const int r = (int)std::floor(2.f * se_size / cell_size_);
    check(gm_scale_uniform(ctx_, r));
    if (r > 1) {
      check(gm_update(ctx_, GM_SE_CUBE));  // as gm_dilate, gm_erode, gm_extract_mesh
      ReadClosingMesh(mesh);
      return;
    }
    // this overload runs the passes at every radius
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
  // closing grid with area-weighted unit normals, then, when bilateral
  // filtering is on, num_bilateral_iters filtering iterations of the normals
  // (GMorpho.cu:6957-6978), copied into mesh.
  void ExtractClosingMesh(Mesh& mesh, int num_bilateral_iters = 20) {
    check(gm_extract_mesh(ctx_));
    ReadClosingMesh(mesh, num_bilateral_iters);
  }
  // The extracted mesh, its normals filtered when bilateral filtering is on,
  // copied into mesh.
  void ReadClosingMesh(Mesh& mesh, int num_bilateral_iters = 20) {
    if (bilateral_filtering_) check(gm_filter_normals(ctx_, num_bilateral_iters));
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
    sf.set_device(ctx_);
    if (sf.mode() != ScaleField::BRUSHED) brushed_id_ = 0;
    switch (sf.mode()) {
      case ScaleField::BRUSHED: {
        // the strokes not yet on the device grid; a new field state starts
        // from the global scale, as UpdateGridGlobalScale does
        if (brushed_id_ != sf.id() || brushed_gen_ != sf.generation() || brushed_n_ > sf.strokes().size()) {
          check(gm_scale_grid_fill(ctx_, std::max(-128, std::min(127, sf.uniform_radius()))));
          brushed_id_ = sf.id();
          brushed_gen_ = sf.generation();
          brushed_n_ = 0;
        }
        for (; brushed_n_ < sf.strokes().size(); ++brushed_n_) {
          const ScaleField::ScalePoint& p = sf.strokes()[brushed_n_];
          const float pos[3] = {p.position()[0], p.position()[1], p.position()[2]};
          check(gm_scale_brush(ctx_, pos, p.scale(), p.support(), sf.global_scale(), (int)p.edit_mode()));
        }
        break;
      }
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
  // brush strokes of which field state are on the context's grid
  unsigned long long brushed_id_ = 0, brushed_gen_ = 0;
  size_t brushed_n_ = 0;
};

}  // namespace MorphoGraphics

#endif
