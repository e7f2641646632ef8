// Synthetic workload generators for the GMorpho closing path (host C++).
//
// These build the seeded input meshes and ScaleField lattices that the
// closing benchmarks and tests run on:
//   * icospheres at subdivision k (10*4^k+2 vertices, 20*4^k faces; k=3 has the
//     642/1280 topology of the only mesh fixture the reference ships,
//     Bin/Resources/Models/sphere.off);
//   * the sphere soup (concatenated icospheres with uniform centres/radii);
//   * the spherical-harmonics blob (icosphere radially displaced by
//     amp * sum_{l<=lmax} a_lm Y_lm, a_lm ~ N(0, 1/(l+1)^2));
//   * the FP32 ScaleField lattice (uniform values, box-smoothed).
// Everything is deterministic: a splitmix64 stream, Box-Muller normals and
// double-precision geometry rounded once to float.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "../../include/gmorpho_meshgen.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0,1)
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  double gauss() {
    double u1 = 1.0 - uniform();  // (0,1]
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct V3 {
  double x, y, z;
};

// Unit icosphere: icosahedron with vertices (0,+-phi,+-1) and cyclic
// permutations, each subdivision splits a face in four and projects the edge
// midpoints back onto the sphere. Faces are counter-clockwise seen from
// outside.
void icosphere(int subdiv, std::vector<V3>& v, std::vector<uint32_t>& f) {
  const double p = (1.0 + std::sqrt(5.0)) / 2.0;
  v = {{0, p, 1},  {0, p, -1},  {0, -p, 1},  {0, -p, -1}, {p, 1, 0},  {p, -1, 0},
       {-p, 1, 0}, {-p, -1, 0}, {1, 0, p},   {-1, 0, p},  {1, 0, -p}, {-1, 0, -p}};
  for (auto& a : v) {
    double n = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    a.x /= n; a.y /= n; a.z /= n;
  }
  // The 20 faces are the vertex triples that are mutually adjacent (squared
  // edge length 4/(1+phi^2) after normalisation), enumerated in lexicographic
  // order and oriented so that (b-a)x(c-a) points outward.
  f.clear();
  auto d2 = [&](int i, int j) {
    double dx = v[i].x - v[j].x, dy = v[i].y - v[j].y, dz = v[i].z - v[j].z;
    return dx * dx + dy * dy + dz * dz;
  };
  const double e2 = 4.0 / (1.0 + p * p);
  auto adj = [&](int i, int j) { return std::fabs(d2(i, j) - e2) < 1e-9; };
  for (int i = 0; i < 12; ++i)
    for (int j = i + 1; j < 12; ++j)
      for (int k = j + 1; k < 12; ++k) {
        if (!adj(i, j) || !adj(j, k) || !adj(i, k)) continue;
        V3 a = v[i], b = v[j], c = v[k];
        V3 e1{b.x - a.x, b.y - a.y, b.z - a.z}, e2v{c.x - a.x, c.y - a.y, c.z - a.z};
        V3 n{e1.y * e2v.z - e1.z * e2v.y, e1.z * e2v.x - e1.x * e2v.z, e1.x * e2v.y - e1.y * e2v.x};
        double s = n.x * (a.x + b.x + c.x) + n.y * (a.y + b.y + c.y) + n.z * (a.z + b.z + c.z);
        f.push_back((uint32_t)i);
        if (s >= 0) { f.push_back((uint32_t)j); f.push_back((uint32_t)k); }
        else        { f.push_back((uint32_t)k); f.push_back((uint32_t)j); }
      }
  for (int s = 0; s < subdiv; ++s) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
      auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      V3 m{(v[a].x + v[b].x) * 0.5, (v[a].y + v[b].y) * 0.5, (v[a].z + v[b].z) * 0.5};
      double n = std::sqrt(m.x * m.x + m.y * m.y + m.z * m.z);
      m.x /= n; m.y /= n; m.z /= n;
      uint32_t id = (uint32_t)v.size();
      v.push_back(m);
      mid.emplace(key, id);
      return id;
    };
    std::vector<uint32_t> nf;
    nf.reserve(f.size() * 4);
    for (size_t i = 0; i < f.size(); i += 3) {
      uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
      uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      uint32_t q[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
      nf.insert(nf.end(), q, q + 12);
    }
    f.swap(nf);
  }
}

// Orthonormal real spherical harmonics Y_l^m at unit direction d
// (associated Legendre functions without the Condon-Shortley phase).
void real_sh(int lmax, const V3& d, std::vector<double>& out) {
  const double PI = 3.141592653589793;
  double ct = d.z, st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
  double phi = std::atan2(d.y, d.x);
  // P[l][m]
  std::vector<double> P((lmax + 1) * (lmax + 1), 0.0);
  auto at = [&](int l, int m) -> double& { return P[l * (lmax + 1) + m]; };
  at(0, 0) = 1.0;
  for (int m = 1; m <= lmax; ++m) at(m, m) = at(m - 1, m - 1) * (2 * m - 1) * st;
  for (int m = 0; m < lmax; ++m) at(m + 1, m) = (2 * m + 1) * ct * at(m, m);
  for (int m = 0; m <= lmax; ++m)
    for (int l = m + 2; l <= lmax; ++l)
      at(l, m) = ((2 * l - 1) * ct * at(l - 1, m) - (l + m - 1) * at(l - 2, m)) / (l - m);
  out.clear();
  for (int l = 0; l <= lmax; ++l) {
    for (int m = -l; m <= l; ++m) {
      int am = m < 0 ? -m : m;
      double ratio = 1.0;  // (l-am)!/(l+am)!
      for (int i = l - am + 1; i <= l + am; ++i) ratio /= i;
      double K = std::sqrt((2 * l + 1) / (4 * PI) * ratio);
      double y;
      if (m > 0) y = std::sqrt(2.0) * K * std::cos(m * phi) * at(l, am);
      else if (m < 0) y = std::sqrt(2.0) * K * std::sin(am * phi) * at(l, am);
      else y = K * at(l, 0);
      out.push_back(y);
    }
  }
}

void write_out(const std::vector<V3>& v, const std::vector<uint32_t>& f, float* P, uint32_t* T) {
  if (P)
    for (size_t i = 0; i < v.size(); ++i) {
      P[3 * i] = (float)v[i].x; P[3 * i + 1] = (float)v[i].y; P[3 * i + 2] = (float)v[i].z;
    }
  if (T) std::memcpy(T, f.data(), f.size() * sizeof(uint32_t));
}

// Uniform scale + translation mapping the bounding box into [0,1]^3 with its
// largest extent equal to 1 (done in double before the float rounding).
void normalize_unit_cube(std::vector<V3>& v) {
  if (v.empty()) return;
  V3 lo = v[0], hi = v[0];
  for (auto& a : v) {
    lo.x = std::fmin(lo.x, a.x); lo.y = std::fmin(lo.y, a.y); lo.z = std::fmin(lo.z, a.z);
    hi.x = std::fmax(hi.x, a.x); hi.y = std::fmax(hi.y, a.y); hi.z = std::fmax(hi.z, a.z);
  }
  double ext = std::fmax(hi.x - lo.x, std::fmax(hi.y - lo.y, hi.z - lo.z));
  if (ext <= 0) return;
  for (auto& a : v) {
    a.x = (a.x - lo.x) / ext; a.y = (a.y - lo.y) / ext; a.z = (a.z - lo.z) / ext;
  }
}

}  // namespace

extern "C" {

int gmg_icosphere_counts(int subdiv, int* nv, int* nf) {
  if (subdiv < 0 || subdiv > 9) return -1;
  long long p = 1;
  for (int i = 0; i < subdiv; ++i) p *= 4;
  *nv = (int)(10 * p + 2);
  *nf = (int)(20 * p);
  return 0;
}

int gmg_icosphere(int subdiv, float* P, uint32_t* T) {
  if (subdiv < 0 || subdiv > 9) return -1;
  std::vector<V3> v;
  std::vector<uint32_t> f;
  icosphere(subdiv, v, f);
  write_out(v, f, P, T);
  return 0;
}

int gmg_sh_blob(int subdiv, uint64_t seed, int lmax, double amp, float* P, uint32_t* T) {
  if (subdiv < 0 || subdiv > 9 || lmax < 0 || lmax > 32) return -1;
  std::vector<V3> v;
  std::vector<uint32_t> f;
  icosphere(subdiv, v, f);
  Rng rng(seed);
  std::vector<double> coef;
  for (int l = 0; l <= lmax; ++l)
    for (int m = -l; m <= l; ++m) coef.push_back(rng.gauss() / (l + 1));
  std::vector<double> y;
  for (auto& a : v) {
    real_sh(lmax, a, y);
    double s = 0.0;
    for (size_t i = 0; i < y.size(); ++i) s += coef[i] * y[i];
    double r = 1.0 + amp * s;
    a.x *= r; a.y *= r; a.z *= r;
  }
  normalize_unit_cube(v);
  write_out(v, f, P, T);
  return 0;
}

int gmg_sphere_soup(int nspheres, int subdiv, uint64_t seed, double cmin, double cmax, double rmin,
                    double rmax, float* P, uint32_t* T) {
  if (nspheres <= 0 || subdiv < 0 || subdiv > 7) return -1;
  std::vector<V3> base, v;
  std::vector<uint32_t> bf, f;
  icosphere(subdiv, base, bf);
  Rng rng(seed);
  for (int s = 0; s < nspheres; ++s) {
    V3 c{rng.uniform(cmin, cmax), rng.uniform(cmin, cmax), rng.uniform(cmin, cmax)};
    double r = rng.uniform(rmin, rmax);
    uint32_t o = (uint32_t)v.size();
    for (auto& a : base) v.push_back({c.x + r * a.x, c.y + r * a.y, c.z + r * a.z});
    for (auto i : bf) f.push_back(o + i);
  }
  normalize_unit_cube(v);
  write_out(v, f, P, T);
  return 0;
}

int gmg_scale_lattice(int L, uint64_t seed, double lo, double hi, int smooth_passes, float* out) {
  if (L < 2 || L > 1024 || smooth_passes < 0) return -1;
  size_t n = (size_t)L * L * L;
  std::vector<double> a(n), b(n);
  Rng rng(seed);
  for (size_t i = 0; i < n; ++i) a[i] = rng.uniform(lo, hi);
  // 3x3x3 box filter with clamp-to-edge borders, repeated.
  for (int pass = 0; pass < smooth_passes; ++pass) {
    for (int z = 0; z < L; ++z)
      for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x) {
          double s = 0.0;
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                int xx = x + dx, yy = y + dy, zz = z + dz;
                xx = xx < 0 ? 0 : (xx >= L ? L - 1 : xx);
                yy = yy < 0 ? 0 : (yy >= L ? L - 1 : yy);
                zz = zz < 0 ? 0 : (zz >= L ? L - 1 : zz);
                s += a[(size_t)xx + (size_t)L * (yy + (size_t)L * zz)];
              }
          b[(size_t)x + (size_t)L * (y + (size_t)L * z)] = s / 27.0;
        }
    a.swap(b);
  }
  for (size_t i = 0; i < n; ++i) out[i] = (float)a[i];
  return 0;
}

}  // extern "C"
