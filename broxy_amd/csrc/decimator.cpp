// Proxy decimation (the reference's Decimator, Decimator/Decimator.cpp): HC
// smoothing, then quadric-error edge collapses with bounded placement.
//
// Per edge (v0, v1) the collapse (CQEMEdgeCollapse, Decimator.cpp:210-541):
//  * placement: minimize Q(x) = Q_v0(x) + Q_v1(x) (sums of unit face-plane
//    quadrics) subject to n_f . (x - p_f) >= 0 for every face f around v0 or
//    v1 (Decimator.cpp:666-700): the new vertex stays on the outer side of
//    the faces it replaces, which keeps the proxy around the shape. Eigen
//    directions of the quadratic part whose eigenvalue is below 1e-3 of the
//    largest get the regularizer lambda |x - m|^2 with lambda = 1/|v0 - v1|^2
//    and m the edge midpoint (Decimator.cpp:566-625). The QP is solved by
//    the dual active-set method of Goldfarb and Idnani (1983);
//  * priority (ComputePriority, Decimator.cpp:385-523): the quadric at the
//    placement (at least 1e-15), over the smallest quality of the moved
//    faces clamped to 0.3; 1e11 when the QP is infeasible or a moved face's
//    normal turns by more than 90 degrees, or when the edge is longer than
//    max_edge_length * min(s(v0), s(v1)) * cell_size (scale-grid values);
//  * execution (Execute, Decimator.cpp:247-265): v0 is removed, v1 moves to
//    the placement and takes v0's quadric.
// The loop pops the cheapest edge, skips stale entries (an end vertex was
// removed or moved since the entry was made; other entries around a moved
// vertex keep their priority, as in the reference's heap) and infeasible
// ones, and stops at the target
// face count, at a priority above the target error, or when the heap is
// empty; the current error is the priority of the last valid collapse taken
// from the heap, executed or not (currMetric, Decimator.cpp:1196-1201).
// Before the collapses, num_hc_iters iterations of HC Laplacian smoothing
// (Vollmer, Mencl and Mueller 1999; alpha = 0, beta = 0.5) move the vertices
// (Decimator.cpp:1149).
// Differences, with their reason: each moved face's normal is compared with
// its own previous normal (the reference compares v1's faces with v0's list,
// Decimator.cpp:325); a collapse must keep the mesh a 2-manifold (link
// condition; the reference leaves that check commented out, 281-285); the
// scale-grid lookup clamps to res - 1 (the reference to res, one past).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <string>
#include <vector>

#include "../../include/gmorpho_c.h"
#include "../../include/gmorpho_decimate.h"

namespace {

using V3 = std::array<double, 3>;

inline void sub(const double* a, const double* b, double* o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
inline double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
inline double norm(const double* a) { return std::sqrt(dot(a, a)); }

// Quadric xT A x + b . x + c (A symmetric: a00 a01 a02 a11 a12 a22).
struct Quadric {
  double A[6] = {0, 0, 0, 0, 0, 0}, b[3] = {0, 0, 0}, c = 0;
  void add_plane(const double* n, double d) {  // (n . x + d)^2, |n| = 1
    A[0] += n[0] * n[0];
    A[1] += n[0] * n[1];
    A[2] += n[0] * n[2];
    A[3] += n[1] * n[1];
    A[4] += n[1] * n[2];
    A[5] += n[2] * n[2];
    for (int k = 0; k < 3; ++k) b[k] += 2 * d * n[k];
    c += d * d;
  }
  void operator+=(const Quadric& o) {
    for (int k = 0; k < 6; ++k) A[k] += o.A[k];
    for (int k = 0; k < 3; ++k) b[k] += o.b[k];
    c += o.c;
  }
  double eval(const double* x) const {
    return A[0] * x[0] * x[0] + 2 * A[1] * x[0] * x[1] + 2 * A[2] * x[0] * x[2] + A[3] * x[1] * x[1] +
           2 * A[4] * x[1] * x[2] + A[5] * x[2] * x[2] + dot(b, x) + c;
  }
  void full(double M[3][3]) const {
    M[0][0] = A[0];
    M[0][1] = M[1][0] = A[1];
    M[0][2] = M[2][0] = A[2];
    M[1][1] = A[3];
    M[1][2] = M[2][1] = A[4];
    M[2][2] = A[5];
  }
};

// Eigen decomposition of a symmetric 3x3 matrix (cyclic Jacobi rotations):
// M = V diag(e) V^T, eigenvector k in column k of V.
void eigen_sym3(const double Min[3][3], double e[3], double V[3][3]) {
  double M[3][3];
  std::memcpy(M, Min, sizeof(M));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = M[0][1] * M[0][1] + M[0][2] * M[0][2] + M[1][2] * M[1][2];
    if (off < 1e-30) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (std::fabs(M[p][q]) < 1e-300) continue;
        const double th = (M[q][q] - M[p][p]) / (2 * M[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; ++k) {  // M = J^T M J
          const double mkp = M[k][p], mkq = M[k][q];
          M[k][p] = c * mkp - s * mkq;
          M[k][q] = s * mkp + c * mkq;
        }
        for (int k = 0; k < 3; ++k) {
          const double mpk = M[p][k], mqk = M[q][k];
          M[p][k] = c * mpk - s * mqk;
          M[q][k] = s * mpk + c * mqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int k = 0; k < 3; ++k) e[k] = M[k][k];
}

bool inv3(const double M[3][3], double I[3][3]) {
  const double d = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                   M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                   M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  if (!(std::fabs(d) > 1e-300)) return false;
  I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / d;
  I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / d;
  I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / d;
  I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / d;
  I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / d;
  I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / d;
  I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / d;
  I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / d;
  I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / d;
  return true;
}

// Inverse of a small symmetric positive definite matrix (n <= 3) by
// Gauss-Jordan; false when singular.
bool inv_small(int n, const double* M, double* I) {
  double a[3][6];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j) a[i][j] = j < n ? M[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (!(std::fabs(a[p][c]) > 1e-14)) return false;
    if (p != c)
      for (int j = 0; j < 2 * n; ++j) std::swap(a[p][j], a[c][j]);
    const double inv = 1 / a[c][c];
    for (int j = 0; j < 2 * n; ++j) a[c][j] *= inv;
    for (int r = 0; r < n; ++r)
      if (r != c) {
        const double f = a[r][c];
        for (int j = 0; j < 2 * n; ++j) a[r][j] -= f * a[c][j];
      }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) I[i * n + j] = a[i][n + j];
  return true;
}

// min 1/2 xT G x + g . x  s.t.  N_i . x >= e_i, G positive definite (3x3),
// by the dual active-set method of Goldfarb and Idnani: start at the
// unconstrained minimum, add the most violated constraint, take the partial
// or full step (dropping an active constraint whose multiplier would turn
// negative), until no constraint is violated. Returns false when infeasible.
bool solve_qp(const double G[3][3], const double* g, const std::vector<V3>& Nc, const std::vector<double>& ec,
              double* x) {
  double Gi[3][3];
  if (!inv3(G, Gi)) return false;
  for (int i = 0; i < 3; ++i) x[i] = -(Gi[i][0] * g[0] + Gi[i][1] * g[1] + Gi[i][2] * g[2]);
  const int m = (int)Nc.size();
  std::vector<int> act;
  std::vector<double> u;
  std::vector<char> in(m, 0);
  const double tol = 1e-12;
  for (int outer = 0; outer < 4 * m + 8; ++outer) {
    int p = -1;
    double worst = 0.0;
    for (int i = 0; i < m; ++i) {
      if (in[i]) continue;
      const double s = dot(Nc[i].data(), x) - ec[i];
      const double scale = 1e-9 * (1.0 + std::fabs(ec[i]));
      if (s < -scale && s < worst) {
        worst = s;
        p = i;
      }
    }
    if (p < 0) return true;
    double up = 0.0;
    for (int inner = 0; inner < 4 * m + 8; ++inner) {
      const int q = (int)act.size();
      // Gi n_p, N_A^T Gi N_A and N_A^T Gi n_p
      double Gn[3];
      for (int i = 0; i < 3; ++i) Gn[i] = dot(Gi[i], Nc[p].data());
      double r[3] = {0, 0, 0}, z[3] = {Gn[0], Gn[1], Gn[2]};
      if (q > 0) {
        double MA[9], MI[9], GN[3][3];
        for (int j = 0; j < q; ++j)
          for (int i = 0; i < 3; ++i) GN[j][i] = dot(Gi[i], Nc[act[j]].data());
        for (int i = 0; i < q; ++i)
          for (int j = 0; j < q; ++j) MA[i * q + j] = dot(Nc[act[i]].data(), GN[j]);
        if (!inv_small(q, MA, MI)) return false;
        double t[3];
        for (int i = 0; i < q; ++i) t[i] = dot(Nc[act[i]].data(), Gn);
        for (int i = 0; i < q; ++i) {
          r[i] = 0;
          for (int j = 0; j < q; ++j) r[i] += MI[i * q + j] * t[j];
        }
        for (int j = 0; j < q; ++j)
          for (int i = 0; i < 3; ++i) z[i] -= GN[j][i] * r[j];
      }
      // partial step (dual feasibility) and full step (constraint p met)
      double t1 = std::numeric_limits<double>::infinity();
      int k = -1;
      for (int j = 0; j < q; ++j)
        if (r[j] > tol && u[j] / r[j] < t1) {
          t1 = u[j] / r[j];
          k = j;
        }
      // with 3 active constraints the primal step space is empty (z = 0 in
      // exact arithmetic): only a dual step is possible
      const double zn = q >= 3 ? 0.0 : dot(z, Nc[p].data());
      const double sp = dot(Nc[p].data(), x) - ec[p];
      const double t2 = zn > tol ? -sp / zn : std::numeric_limits<double>::infinity();
      const double t = std::min(t1, t2);
      if (!std::isfinite(t)) return false;  // infeasible
      for (int i = 0; i < 3; ++i) x[i] += std::isfinite(t2) ? t * z[i] : 0.0;
      for (int j = 0; j < q; ++j) u[j] -= t * r[j];
      up += t;
      if (t2 <= t1) {
        act.push_back(p);
        u.push_back(up);
        in[p] = 1;
        break;
      }
      in[act[k]] = 0;
      act.erase(act.begin() + k);
      u.erase(u.begin() + k);
    }
  }
  return false;
}

struct Entry {
  double pri;
  int v0, v1;
  uint32_t m0, m1;
  float x[3];
  bool operator<(const Entry& o) const { return pri > o.pri; }  // min-heap
};

}  // namespace

struct gmd_session {
  std::string err;
  std::vector<double> P;  // 3 per vertex
  std::vector<int> F;     // 3 per face, -1 when removed
  std::vector<std::vector<int>> vf;  // faces around each vertex
  std::vector<Quadric> Q;
  std::vector<uint32_t> mark;
  std::vector<char> alive;
  std::priority_queue<Entry> heap;
  int nfaces = 0, nverts = 0, target = 0;
  float targetError = 0.f, metric = 0.f;
  bool constrained = true;
  float alpha = 0.f, cell = 1.f;
  float bmin[3] = {0, 0, 0};
  int res[3] = {0, 0, 0};
  std::vector<int8_t> scale;

  const double* p(int v) const { return &P[3 * (size_t)v]; }
  void face_normal(int f, double* n) const {
    double e1[3], e2[3];
    sub(p(F[3 * f + 1]), p(F[3 * f]), e1);
    sub(p(F[3 * f + 2]), p(F[3 * f]), e2);
    cross(e1, e2, n);
    const double l = norm(n);
    if (l > 0)
      for (int k = 0; k < 3; ++k) n[k] /= l;
  }
  // 4 sqrt(3) area / sum of squared edges: 1 for an equilateral triangle
  static double quality(const double* a, const double* b, const double* c) {
    double e1[3], e2[3], e3[3], n[3];
    sub(b, a, e1);
    sub(c, a, e2);
    sub(c, b, e3);
    cross(e1, e2, n);
    const double s = dot(e1, e1) + dot(e2, e2) + dot(e3, e3);
    return s > 0 ? 2 * std::sqrt(3.0) * norm(n) / s : 0.0;
  }
  double scale_at(const double* x) const {
    if (scale.empty()) return std::numeric_limits<double>::infinity();
    int id[3];
    for (int a = 0; a < 3; ++a) {
      const double g = std::floor((x[a] - bmin[a]) / cell);
      id[a] = (int)std::min(std::max(g, 0.0), (double)(res[a] - 1));
    }
    return scale[(size_t)id[0] + (size_t)res[0] * ((size_t)id[1] + (size_t)res[1] * id[2])];
  }
  bool has(int f, int v) const { return F[3 * f] == v || F[3 * f + 1] == v || F[3 * f + 2] == v; }

  void neighbours(int v, std::vector<int>& out) const {
    out.clear();
    for (int f : vf[v])
      for (int k = 0; k < 3; ++k)
        if (F[3 * f + k] != v) out.push_back(F[3 * f + k]);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }

  // Priority and placement of collapsing v0 into v1 (1e11: never).
  double evaluate(int v0, int v1, double* x, bool* feasible) const {
    *feasible = false;
    Quadric q = Q[v0];
    q += Q[v1];
    double A[3][3], e[3], Vv[3][3];
    q.full(A);
    eigen_sym3(A, e, Vv);
    const double emax = std::max(std::fabs(e[0]), std::max(std::fabs(e[1]), std::fabs(e[2])));
    double mid[3], ed[3];
    for (int k = 0; k < 3; ++k) mid[k] = 0.5 * (p(v0)[k] + p(v1)[k]);
    sub(p(v0), p(v1), ed);
    const double len2 = dot(ed, ed);
    const double lambda = len2 > 0 ? 1.0 / len2 : 1.0;
    // objective xT A x + b . x (+ c) with the regularizer, as 1/2 xT G x + g . x
    double G[3][3], g[3] = {q.b[0], q.b[1], q.b[2]};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) G[i][j] = 2 * A[i][j];
    for (int k = 0; k < 3; ++k)
      if (emax > 0 ? std::fabs(e[k]) / emax < 1e-3 : true) {
        const double w[3] = {Vv[0][k], Vv[1][k], Vv[2][k]};
        const double wm = dot(w, mid);
        for (int i = 0; i < 3; ++i) {
          for (int j = 0; j < 3; ++j) G[i][j] += 2 * lambda * w[i] * w[j];
          g[i] -= 2 * lambda * wm * w[i];
        }
      }
    std::vector<V3> Nc;
    std::vector<double> ec;
    if (constrained) {
      // faces around v0 without v1, then every face around v1
      // (Decimator.cpp:666-700)
      for (int pass = 0; pass < 2; ++pass) {
        const int v = pass == 0 ? v0 : v1;
        for (int f : vf[v]) {
          if (pass == 0 && has(f, v1)) continue;
          double n[3];
          face_normal(f, n);
          Nc.push_back(V3{n[0], n[1], n[2]});
          ec.push_back(dot(n, p(F[3 * f])));
        }
      }
    }
    if (!solve_qp(G, g, Nc, ec, x)) return 1e11;
    *feasible = true;
    double obj = q.eval(x);
    if (obj < 1e-15) obj = 1e-15;
    // moved faces: normal variation and quality
    double minCos = 1e100, minQual = 1e100;
    for (int pass = 0; pass < 2; ++pass) {
      const int v = pass == 0 ? v0 : v1, o = pass == 0 ? v1 : v0;
      for (int f : vf[v]) {
        if (has(f, o)) continue;
        double n0[3], n1[3], pts[3][3];
        face_normal(f, n0);
        for (int k = 0; k < 3; ++k) {
          const int w = F[3 * f + k];
          for (int a = 0; a < 3; ++a) pts[k][a] = w == v ? x[a] : p(w)[a];
        }
        double e1[3], e2[3];
        sub(pts[1], pts[0], e1);
        sub(pts[2], pts[0], e2);
        cross(e1, e2, n1);
        const double l = norm(n1);
        const double c = l > 0 ? dot(n0, n1) / l : -1.0;
        minCos = std::min(minCos, c);
        minQual = std::min(minQual, quality(pts[0], pts[1], pts[2]));
      }
    }
    if (minQual > 0.3) minQual = 0.3;
    if (minCos < 0) return 1e11;
    double pri = minQual > 0 ? obj / minQual : 1e11;
    const double elen = std::sqrt(len2);
    if (elen > alpha * std::min(scale_at(p(v0)), scale_at(p(v1))) * cell) pri = 1e11;
    return pri;
  }

  // Collapsing v0 into v1 keeps a closed 2-manifold: the common neighbours
  // are exactly the third vertices of the two faces on the edge.
  bool link_ok(int v0, int v1) const {
    std::vector<int> a, b;
    neighbours(v0, a);
    neighbours(v1, b);
    std::vector<int> common;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(common));
    int shared = 0;
    for (int f : vf[v0])
      if (has(f, v1)) ++shared;
    if (shared != 2 || (int)common.size() != 2) return false;
    return nverts > 4;
  }

  void push(int v0, int v1) {
    double x[3] = {0, 0, 0};
    bool ok;
    const double pri = evaluate(v0, v1, x, &ok);
    heap.push(Entry{ok ? pri : 1e11, v0, v1, mark[v0], mark[v1], {(float)x[0], (float)x[1], (float)x[2]}});
  }

  void collapse(int v0, int v1, const float* xf) {
    for (int k = 0; k < 3; ++k) P[3 * (size_t)v1 + k] = xf[k];
    Q[v1] += Q[v0];
    std::vector<int> keep;
    for (int f : vf[v1])
      if (!has(f, v0)) keep.push_back(f);
    for (int f : vf[v0]) {
      if (has(f, v1)) {
        F[3 * f] = F[3 * f + 1] = F[3 * f + 2] = -1;
        --nfaces;
        continue;
      }
      for (int k = 0; k < 3; ++k)
        if (F[3 * f + k] == v0) F[3 * f + k] = v1;
      keep.push_back(f);
    }
    // the removed faces leave the lists of their third vertices too
    std::vector<int> nb;
    for (int f : keep)
      for (int k = 0; k < 3; ++k) nb.push_back(F[3 * f + k]);
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    for (int w : nb) {
      if (w == v1) continue;
      auto& l = vf[w];
      l.erase(std::remove_if(l.begin(), l.end(), [&](int f) { return F[3 * f] < 0; }), l.end());
    }
    vf[v1] = keep;
    vf[v0].clear();
    alive[v0] = 0;
    --nverts;
    ++mark[v0];
    ++mark[v1];
    for (int w : nb)
      if (w != v1) push(std::min(w, v1), std::max(w, v1));
  }
};

namespace {
// HC Laplacian smoothing (Vollmer, Mencl and Mueller 1999), alpha = 0,
// beta = 0.5, uniform weights over the edge neighbours.
void hc_smooth(gmd_session* s, int iters) {
  const int nv = (int)s->alive.size();
  std::vector<std::vector<int>> nb(nv);
  for (int v = 0; v < nv; ++v) s->neighbours(v, nb[v]);
  std::vector<double> q, pp(3 * (size_t)nv), bb(3 * (size_t)nv);
  for (int it = 0; it < iters; ++it) {
    q = s->P;
    for (int v = 0; v < nv; ++v) {
      for (int a = 0; a < 3; ++a) {
        double acc = 0;
        for (int w : nb[v]) acc += q[3 * (size_t)w + a];
        pp[3 * (size_t)v + a] = nb[v].empty() ? q[3 * (size_t)v + a] : acc / nb[v].size();
        bb[3 * (size_t)v + a] = pp[3 * (size_t)v + a] - q[3 * (size_t)v + a];  // alpha = 0: b = p - q
      }
    }
    for (int v = 0; v < nv; ++v)
      for (int a = 0; a < 3; ++a) {
        double acc = 0;
        for (int w : nb[v]) acc += bb[3 * (size_t)w + a];
        const double d = 0.5 * bb[3 * (size_t)v + a] + (nb[v].empty() ? 0.0 : 0.5 * acc / nb[v].size());
        s->P[3 * (size_t)v + a] = pp[3 * (size_t)v + a] - d;
      }
  }
}
}  // namespace

extern "C" {

int gmd_create(const float* P, int nv, const uint32_t* T, int nt, int target_num_faces, float target_error,
               int use_linear_constraints, float max_edge_length, const float bbox_min[3], const int res[3],
               float cell_size, const int8_t* scale_grid, int num_hc_iters, gmd_session** out) {
  if (!out) return GM_ERR_ARG;
  *out = nullptr;
  if (!P || !T || nv <= 0 || nt <= 0) return GM_ERR_ARG;
  for (size_t i = 0; i < 3 * (size_t)nt; ++i)
    if (T[i] >= (uint32_t)nv) return GM_ERR_ARG;
  if (scale_grid && (!bbox_min || !res || res[0] <= 0 || res[1] <= 0 || res[2] <= 0 || !(cell_size > 0)))
    return GM_ERR_ARG;
  gmd_session* s = new gmd_session();
  s->P.assign(P, P + 3 * (size_t)nv);
  s->F.assign(T, T + 3 * (size_t)nt);
  s->vf.assign(nv, {});
  for (int f = 0; f < nt; ++f)
    for (int k = 0; k < 3; ++k) s->vf[s->F[3 * f + k]].push_back(f);
  s->alive.assign(nv, 1);
  s->mark.assign(nv, 0);
  s->nfaces = nt;
  s->nverts = nv;
  for (int v = 0; v < nv; ++v)
    if (s->vf[v].empty()) {
      s->alive[v] = 0;
      --s->nverts;
    }
  s->target = target_num_faces;
  s->targetError = target_error;
  s->constrained = use_linear_constraints != 0;
  s->alpha = max_edge_length;
  s->cell = cell_size > 0 ? cell_size : 1.f;
  if (scale_grid) {
    for (int a = 0; a < 3; ++a) {
      s->bmin[a] = bbox_min[a];
      s->res[a] = res[a];
    }
    s->scale.assign(scale_grid, scale_grid + (size_t)res[0] * res[1] * res[2]);
  }
  if (num_hc_iters > 0) hc_smooth(s, num_hc_iters);
  s->Q.assign(nv, Quadric());
  for (int f = 0; f < nt; ++f) {
    double n[3];
    s->face_normal(f, n);
    const double d = -dot(n, s->p(s->F[3 * f]));
    for (int k = 0; k < 3; ++k) s->Q[s->F[3 * f + k]].add_plane(n, d);
  }
  for (int f = 0; f < nt; ++f)
    for (int k = 0; k < 3; ++k) {
      const int a = s->F[3 * f + k], b = s->F[3 * f + (k + 1) % 3];
      if (a < b) s->push(a, b);  // each edge once (from one of its two faces)
    }
  *out = s;
  return GM_OK;
}

int gmd_destroy(gmd_session* s) {
  delete s;
  return GM_OK;
}

const char* gmd_last_error(const gmd_session* s) { return s ? s->err.c_str() : "Decimator Error: null session"; }

int gmd_optimize(gmd_session* s, double budget, int* heap_not_empty) {
  if (!s) return GM_ERR_ARG;
  const auto t0 = std::chrono::steady_clock::now();
  while (!s->heap.empty() && s->nfaces > s->target) {
    if (budget > 0 &&
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > budget)
      break;
    const Entry e = s->heap.top();
    s->heap.pop();
    if (!s->alive[e.v0] || !s->alive[e.v1] || e.m0 != s->mark[e.v0] || e.m1 != s->mark[e.v1]) continue;
    if (e.pri >= 1e11 || !s->link_ok(e.v0, e.v1)) continue;
    // the current metric is the cheapest valid collapse's priority; above
    // the target error the optimization stops (and the entry stays queued)
    s->metric = (float)e.pri;
    if (e.pri > s->targetError) {
      s->heap.push(e);
      break;
    }
    s->collapse(e.v0, e.v1, e.x);
  }
  if (heap_not_empty) *heap_not_empty = s->heap.empty() ? 0 : 1;
  return GM_OK;
}

int gmd_state(const gmd_session* s, int* num_faces, int* num_vertices, float* error) {
  if (!s) return GM_ERR_ARG;
  if (num_faces) *num_faces = s->nfaces;
  if (num_vertices) *num_vertices = s->nverts;
  if (error) *error = s->metric;
  return GM_OK;
}

int gmd_get_mesh(const gmd_session* s, float* P, uint32_t* T) {
  if (!s || !P || !T) return GM_ERR_ARG;
  std::vector<int> idx(s->alive.size(), -1);
  int n = 0;
  for (size_t v = 0; v < s->alive.size(); ++v)
    if (s->alive[v]) {
      idx[v] = n;
      for (int a = 0; a < 3; ++a) P[3 * (size_t)n + a] = (float)s->P[3 * v + a];
      ++n;
    }
  size_t t = 0;
  for (size_t f = 0; f * 3 < s->F.size(); ++f) {
    if (s->F[3 * f] < 0) continue;
    for (int k = 0; k < 3; ++k) T[3 * t + k] = (uint32_t)idx[s->F[3 * f + k]];
    ++t;
  }
  return GM_OK;
}

}  // extern "C"
