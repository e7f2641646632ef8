// GMorpho oracle: CPU restatement of the reference's closing path.
//
// TEST INFRASTRUCTURE ONLY -- the checker for the HIP library and the timed CPU
// baseline. It is never linked into, loaded by or called from the product path.
//
// Every function follows the reference's per-voxel logic (file:line cited per
// function) with the float tests replaced by equivalent integer predicates:
// all query points are integers or half-integers and every node size is a
// power of two, so the reference's single-precision distances are exact at the
// leaves and inner-node pruning never removes a passing leaf.
#include "gmorpho_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

typedef long long i64;
typedef unsigned long long u64;

inline int offx(int k) { return k & 1; }
inline int offy(int k) { return (k >> 1) & 1; }
inline int offz(int k) { return (k >> 2) & 1; }

inline size_t cidx(int R, int x, int y, int z) { return (size_t)x + (size_t)R * ((size_t)y + (size_t)R * z); }

inline bool fine(const uint8_t* g, int R, int x, int y, int z) {
  return (g[cidx(R, x >> 1, y >> 1, z >> 1)] >> ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2))) & 1;
}

// Radius of fine voxel p as the max-mipmap level 0 stores it
// (Grid.cu:990-992 for a ScaleField grid, Grid.cu:725 for a uniform size).
struct Radius {
  const int8_t* scale;
  int r;
  int N;
  int at(int x, int y, int z) const {
    if (!scale) return r & 0x7f;
    int s = scale[(size_t)x + (size_t)N * ((size_t)y + (size_t)N * z)];
    return s < 2 ? 2 : (s > 127 ? 127 : s);
  }
};

// Morton code with 10 bits per axis, x in the lowest bit (cuda_math.h:98-153).
inline uint32_t part1by2(uint32_t x) {
  x &= 0x3ff;
  uint32_t r = 0;
  for (int i = 0; i < 10; ++i) r |= ((x >> i) & 1u) << (3 * i);
  return r;
}
inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return (part1by2(z) << 2) | (part1by2(y) << 1) | part1by2(x);
}

// ---------------------------------------------------------------------------
// Max-mipmap (Grid.cu:957-998 level 0, Grid.cu:733-805 upper levels).
// Texel bits 7k..7k+6: radius of child k (max over its subtree); bit 56+k: child
// k holds any occupied fine voxel. Level 0 texel = one cell, children = its 8
// fine voxels. Grids whose R is not a power of two are padded with empty cells.
struct Pyramid {
  int P = 0;       // padded cells per axis (power of two)
  int levels = 0;  // log2(P) + 1
  std::vector<std::vector<u64>> lv;
  u64 at(int L, int x, int y, int z) const {
    int s = P >> L;
    return lv[L][(size_t)x + (size_t)s * ((size_t)y + (size_t)s * z)];
  }
};

Pyramid build_pyramid(const uint8_t* cells, int R, const Radius& rad) {
  Pyramid py;
  py.P = 1;
  while (py.P < R) py.P <<= 1;
  py.levels = 1;
  while ((1 << (py.levels - 1)) < py.P) py.levels++;
  int P = py.P;
  py.lv.resize(py.levels);
  py.lv[0].assign((size_t)P * P * P, 0);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < R; ++z)
    for (int y = 0; y < R; ++y)
      for (int x = 0; x < R; ++x) {
        u64 t = (u64)cells[cidx(R, x, y, z)] << 56;
        for (int k = 0; k < 8; ++k)
          t |= (u64)(rad.at(2 * x + offx(k), 2 * y + offy(k), 2 * z + offz(k)) & 0x7f) << (7 * k);
        py.lv[0][(size_t)x + (size_t)P * ((size_t)y + (size_t)P * z)] = t;
      }
  for (int L = 1; L < py.levels; ++L) {
    int s = P >> L;
    py.lv[L].assign((size_t)s * s * s, 0);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < s; ++z)
      for (int y = 0; y < s; ++y)
        for (int x = 0; x < s; ++x) {
          u64 t = 0;
          for (int l = 0; l < 8; ++l) {
            u64 c = py.at(L - 1, 2 * x + offx(l), 2 * y + offy(l), 2 * z + offz(l));
            int m = 0;
            for (int k = 0; k < 8; ++k) m = std::max(m, (int)((c >> (7 * k)) & 0x7f));
            t |= (u64)m << (7 * l);
            if ((c >> 56) != 0) t |= 1ull << (56 + l);
          }
          py.lv[L][(size_t)x + (size_t)s * ((size_t)y + (size_t)s * z)] = t;
        }
  }
  return py;
}

// Squared distance along one axis from point q (in some unit) to [lo, hi].
inline i64 axd(i64 q, i64 lo, i64 hi) {
  i64 d = q < lo ? lo - q : (q > hi ? q - hi : 0);
  return d * d;
}

struct Node {
  int d, x, y, z;
};

// Dilation query (GMorpho.cu:2860-2925, 2995-3037): does a fine voxel q
// occupied in the pyramid satisfy SqDist(P, [q,q+1]^3) < rho^2 (sphere), or
// [q,q+1]^3 touch the closed box [P-rho, P+rho]^3 (cube)? P is a lattice point
// in fine units.
bool dfs_dilate(const Pyramid& py, int Px, int Py, int Pz, int rho, int se) {
  const int maxd = py.levels;
  const i64 r2 = (i64)rho * rho;
  std::vector<Node> st;
  st.reserve(256);
  st.push_back({0, 0, 0, 0});
  while (!st.empty()) {
    Node n = st.back();
    st.pop_back();
    u64 t = py.at(maxd - 1 - n.d, n.x, n.y, n.z);
    int occ = (int)(t >> 56);
    i64 cs = (i64)(2 * py.P) >> (n.d + 1);
    for (int k = 0; k < 8; ++k) {
      if (!((occ >> k) & 1)) continue;
      int cx = 2 * n.x + offx(k), cy = 2 * n.y + offy(k), cz = 2 * n.z + offz(k);
      i64 lx = cx * cs, ly = cy * cs, lz = cz * cs;
      bool hit;
      if (se == 0) {
        hit = axd(Px, lx, lx + cs) + axd(Py, ly, ly + cs) + axd(Pz, lz, lz + cs) < r2;
      } else {
        hit = (Px - rho <= lx + cs && Px + rho >= lx) && (Py - rho <= ly + cs && Py + rho >= ly) &&
              (Pz - rho <= lz + cs && Pz + rho >= lz);
      }
      if (!hit) continue;
      if (n.d + 1 == maxd) return true;
      st.push_back({n.d + 1, cx, cy, cz});
    }
  }
  return false;
}

// Erosion query (GMorpho.cu:4376-4645, 5081-5345): does a fine voxel q of the
// contour satisfy SqDist(Q, [q,q+1]^3) < (r(q)-1)^2, where inner nodes use the
// subtree's max radius? Q is given doubled (Q2 = 2Q) so half-integer queries
// stay integral; distances are then in quarter units.
bool dfs_erode(const Pyramid& py, int Qx2, int Qy2, int Qz2) {
  const int maxd = py.levels;
  std::vector<Node> st;
  st.reserve(256);
  st.push_back({0, 0, 0, 0});
  while (!st.empty()) {
    Node n = st.back();
    st.pop_back();
    u64 t = py.at(maxd - 1 - n.d, n.x, n.y, n.z);
    int occ = (int)(t >> 56);
    i64 cs2 = (i64)(4 * py.P) >> (n.d + 1);  // child size in half units
    for (int k = 0; k < 8; ++k) {
      if (!((occ >> k) & 1)) continue;
      int rr = (int)((t >> (7 * k)) & 0x7f) - 1;
      int cx = 2 * n.x + offx(k), cy = 2 * n.y + offy(k), cz = 2 * n.z + offz(k);
      i64 lx = cx * cs2, ly = cy * cs2, lz = cz * cs2;
      i64 d2 = axd(Qx2, lx, lx + cs2) + axd(Qy2, ly, ly + cs2) + axd(Qz2, lz, lz + cs2);
      if (!(d2 < 4 * (i64)rr * rr)) continue;
      if (n.d + 1 == maxd) return true;
      st.push_back({n.d + 1, cx, cy, cz});
    }
  }
  return false;
}

// Exhaustive versions of the two predicates over a bounded window.
bool brute_dilate(const uint8_t* I, int R, int Px, int Py, int Pz, int rho, int se) {
  int N = 2 * R;
  const i64 r2 = (i64)rho * rho;
  int x0 = std::max(0, Px - rho - 1), x1 = std::min(N - 1, Px + rho);
  int y0 = std::max(0, Py - rho - 1), y1 = std::min(N - 1, Py + rho);
  int z0 = std::max(0, Pz - rho - 1), z1 = std::min(N - 1, Pz + rho);
  for (int z = z0; z <= z1; ++z)
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        if (!fine(I, R, x, y, z)) continue;
        if (se != 0) return true;  // the window is exactly the cube
        if (axd(Px, x, x + 1) + axd(Py, y, y + 1) + axd(Pz, z, z + 1) < r2) return true;
      }
  return false;
}

bool brute_erode(const uint8_t* K, int R, const Radius& rad, int rmax, int Qx2, int Qy2, int Qz2) {
  int N = 2 * R;
  int w = rmax + 1;
  int cx = Qx2 / 2, cy = Qy2 / 2, cz = Qz2 / 2;
  for (int z = std::max(0, cz - w); z <= std::min(N - 1, cz + w); ++z)
    for (int y = std::max(0, cy - w); y <= std::min(N - 1, cy + w); ++y)
      for (int x = std::max(0, cx - w); x <= std::min(N - 1, cx + w); ++x) {
        if (!fine(K, R, x, y, z)) continue;
        int rr = rad.at(x, y, z) - 1;
        i64 d2 = axd(Qx2, 2 * x, 2 * x + 2) + axd(Qy2, 2 * y, 2 * y + 2) + axd(Qz2, 2 * z, 2 * z + 2);
        if (d2 < 4 * (i64)rr * rr) return true;
      }
  return false;
}

// Tag predicates of the contour extractions.
// Grid.cu:1572-1607: cell c in [0,R-2]^3 whose 2x2x2 block mixes full (0xff)
// and non-full cells.
bool tag_tight(const uint8_t* G, int R, int x, int y, int z) {
  if (x >= R - 1 || y >= R - 1 || z >= R - 1) return false;
  int ind = 0;
  for (int k = 0; k < 8; ++k)
    if (G[cidx(R, x + offx(k), y + offy(k), z + offz(k))] == 0xff) ind |= 1 << k;
  return ind != 0 && ind != 0xff;
}
// Grid.cu:1499-1570: cell c in [1,R-2]^3 with a 6-neighbour byte != its byte.
bool tag_base(const uint8_t* G, int R, int x, int y, int z) {
  if (x < 1 || y < 1 || z < 1 || x >= R - 1 || y >= R - 1 || z >= R - 1) return false;
  uint8_t c = G[cidx(R, x, y, z)];
  return G[cidx(R, x + 1, y, z)] != c || G[cidx(R, x - 1, y, z)] != c ||
         G[cidx(R, x, y + 1, z)] != c || G[cidx(R, x, y - 1, z)] != c ||
         G[cidx(R, x, y, z + 1)] != c || G[cidx(R, x, y, z - 1)] != c;
}

// Marching-cubes corner order (MarchingCubesMesher.h:39-46 with the corner
// assignment of GMorpho.cu:6156-6242): corner k of the cube whose min fine
// voxel is f sits at f + kCorner[k].
const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

inline int mc_case(const uint8_t* C, int R, int fx, int fy, int fz) {
  int v = 0;
  for (int k = 0; k < 8; ++k)
    if (fine(C, R, fx + kCorner[k][0], fy + kCorner[k][1], fz + kCorner[k][2])) v |= 1 << k;
  return v;
}
inline bool mixed(int v) { return v != 0 && v != 255; }

// Grid.cu:1329-1462: cell c in [1,R-2]^3 one of whose 8 fine cubes is mixed.
bool tag_mcm(const uint8_t* G, int R, int x, int y, int z) {
  if (x < 1 || y < 1 || z < 1 || x >= R - 1 || y >= R - 1 || z >= R - 1) return false;
  for (int j = 0; j < 8; ++j)
    if (mixed(mc_case(G, R, 2 * x + offx(j), 2 * y + offy(j), 2 * z + offz(j)))) return true;
  return false;
}

std::vector<uint32_t> contour_list(const uint8_t* G, int R, int which) {
  std::vector<std::vector<uint32_t>> part(R);
#pragma omp parallel for schedule(dynamic)
  for (int z = 0; z < R; ++z)
    for (int y = 0; y < R; ++y)
      for (int x = 0; x < R; ++x) {
        bool t = which == 0 ? tag_tight(G, R, x, y, z)
                            : (which == 1 ? tag_base(G, R, x, y, z) : tag_mcm(G, R, x, y, z));
        if (t) part[z].push_back((uint32_t)cidx(R, x, y, z));
      }
  std::vector<uint32_t> out;
  for (auto& p : part) out.insert(out.end(), p.begin(), p.end());
  // thrust::sort_by_key on EncodeMorton3 (Grid.cu:1939-1958); codes are unique.
  std::vector<std::pair<uint32_t, uint32_t>> kv(out.size());
  for (size_t i = 0; i < out.size(); ++i) {
    uint32_t c = out[i];
    uint32_t x = c % R, y = (c / R) % R, z = c / ((uint32_t)R * R);
    kv[i] = {morton3(x, y, z), c};
  }
  std::sort(kv.begin(), kv.end());
  for (size_t i = 0; i < out.size(); ++i) out[i] = kv[i].second;
  return out;
}

int dR_axis(const int* dr, int a, int R) {
  if (!dr) return R;
  int v = dr[a];
  return v < 0 ? 0 : (v > R ? R : v);
}

// ---------------------------------------------------------------------------
// Dilation (DilateBySphereMipmap, GMorpho.cu:3538-3596; cube: 3231-3314).
void dilate_cells(const uint8_t* I, int R, const int* dr, const Radius& rad, int se, int brute,
                  const Pyramid* py, uint8_t* D, int zlo, int zhi, bool fine_all) {
  auto pred = [&](int px, int py_, int pz) {
    int rr = rad.at(px, py_, pz);
    int rho = se == 0 ? rr + 1 : rr;
    return brute ? brute_dilate(I, R, px, py_, pz, rho, se) : dfs_dilate(*py, px, py_, pz, rho, se);
  };
  int dx = dR_axis(dr, 0, R), dy = dR_axis(dr, 1, R), dz = dR_axis(dr, 2, R);
  // Coarse pass (GMorpho.cu:2828-2928 / 1838-1934): one query at the cell's
  // min corner with sub-voxel 0's radius; a hit fills the whole cell.
#pragma omp parallel for schedule(dynamic)
  for (int z = zlo; z < std::min(zhi, dz); ++z)
    for (int y = 0; y < dy; ++y)
      for (int x = 0; x < dx; ++x) {
        size_t i = cidx(R, x, y, z);
        if (I[i] == 0xff) continue;
        if (pred(2 * x, 2 * y, 2 * z)) D[i] = 0xff;
      }
  if (!fine_all) return;
  // Tight contour of the coarse result, then the fine pass: 8 queries per
  // contour cell overwrite the cell (GMorpho.cu:2930-3056 / 2603-2721).
  std::vector<uint32_t> T = contour_list(D, R, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long t = 0; t < (long long)T.size(); ++t) {
    uint32_t c = T[t];
    int x = c % R, y = (c / R) % R, z = c / (R * R);
    if (z < zlo || z >= zhi) continue;
    uint8_t b = 0;
    for (int k = 0; k < 8; ++k)
      if (pred(2 * x + offx(k), 2 * y + offy(k), 2 * z + offz(k))) b |= 1 << k;
    D[c] = b;
  }
}

void contour_K(const uint8_t* D, int R, uint8_t* K) {
  // Grid.cu:1230-1327: empty fine voxels with an occupied 6-neighbour, for
  // cells in [1,R-2]^3 that are not full; every other cell is 0.
  std::memset(K, 0, (size_t)R * R * R);
#pragma omp parallel for schedule(static)
  for (int z = 1; z < R - 1; ++z)
    for (int y = 1; y < R - 1; ++y)
      for (int x = 1; x < R - 1; ++x) {
        uint8_t c = D[cidx(R, x, y, z)];
        if (c == 0xff) continue;
        uint8_t tag = 0;
        for (int k = 0; k < 8; ++k) {
          int fx = 2 * x + offx(k), fy = 2 * y + offy(k), fz = 2 * z + offz(k);
          bool n = fine(D, R, fx + 1, fy, fz) || fine(D, R, fx - 1, fy, fz) ||
                   fine(D, R, fx, fy + 1, fz) || fine(D, R, fx, fy - 1, fz) ||
                   fine(D, R, fx, fy, fz + 1) || fine(D, R, fx, fy, fz - 1);
          if (n) tag |= 1 << k;
        }
        K[cidx(R, x, y, z)] = tag & (uint8_t)~c;
      }
}

void erode_cells(const uint8_t* I, const uint8_t* D, const uint8_t* K, int R, const int* dr,
                 const Radius& rad, int brute, const Pyramid* pk, uint8_t* C, int zlo, int zhi,
                 bool fine_all) {
  int rmax = 0;
  if (brute) {
    if (!rad.scale) rmax = rad.r & 0x7f;
    else for (size_t i = 0; i < (size_t)8 * R * R * R; ++i) {
      int s = rad.scale[i];
      rmax = std::max(rmax, s < 2 ? 2 : (s > 127 ? 127 : s));
    }
  }
  auto pred = [&](int qx2, int qy2, int qz2) {
    return brute ? brute_erode(K, R, rad, rmax, qx2, qy2, qz2) : dfs_erode(*pk, qx2, qy2, qz2);
  };
  int dx = dR_axis(dr, 0, R), dy = dR_axis(dr, 1, R), dz = dR_axis(dr, 2, R);
  // Coarse pass (GMorpho.cu:4349-4648): query the cell centre 2c+1.
#pragma omp parallel for schedule(dynamic)
  for (int z = zlo; z < std::min(zhi, dz); ++z)
    for (int y = 0; y < dy; ++y)
      for (int x = 0; x < dx; ++x) {
        size_t i = cidx(R, x, y, z);
        if (D[i] == 0 || I[i] == 0xff) continue;
        if (pred(2 * (2 * x + 1), 2 * (2 * y + 1), 2 * (2 * z + 1))) C[i] = 0;
      }
  if (!fine_all) return;
  // 6-connected base contour of the coarse result, then 8 queries at the
  // sub-voxel centres; the result replaces the cell (GMorpho.cu:5041-5374).
  std::vector<uint32_t> U = contour_list(C, R, 1);
  std::vector<uint8_t> out(U.size());
#pragma omp parallel for schedule(dynamic, 64)
  for (long long t = 0; t < (long long)U.size(); ++t) {
    uint32_t c = U[t];
    int x = c % R, y = (c / R) % R, z = c / (R * R);
    uint8_t b = 0xff;
    if (z >= zlo && z < zhi)
      for (int k = 0; k < 8; ++k)
        if (pred(2 * (2 * x + offx(k)) + 1, 2 * (2 * y + offy(k)) + 1, 2 * (2 * z + offz(k)) + 1))
          b &= (uint8_t) ~(1 << k);
    out[t] = b;
  }
  for (size_t t = 0; t < U.size(); ++t) {
    uint32_t c = U[t];
    int z = c / (R * R);
    if (z >= zlo && z < zhi) C[c] = out[t];
  }
}

// ---------------------------------------------------------------------------
// Marching-cubes case tables. The edge mask of a case is derived from the
// corner signs (edge e joins corners kEdge[e]). The triangulation table is
// Bourke/Bloyd's public-domain table; it is stored as one hex string per case
// (edge ids, three per triangle) and must equal the reference's tables.h
// (GMorpho/tables.h:321-578), which is the same public table.
const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                          {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

#include "mc_tritable.inc"

// The reference counts triangles with its own per-case table
// (tables.h:61-318) and emits every triangle of the triangulation table
// (MarchingCubesMesher.cu:1486-1493). The two agree except for case 91, which
// is counted as 4 triangles but triangulated with 5: its 5th triangle lands in
// the first slot of the next cube (or past the end of the list).
struct McTables {
  uint16_t edge[256];
  uint8_t ntri[256];  // count used for allocation and scans
  uint8_t len[256];   // triangles actually listed
  int8_t tri[256][16];
  McTables() {
    for (int c = 0; c < 256; ++c) {
      edge[c] = 0;
      for (int e = 0; e < 12; ++e)
        if (((c >> kEdge[e][0]) & 1) != ((c >> kEdge[e][1]) & 1)) edge[c] |= 1 << e;
      const char* s = kTriHex[c];
      int n = (int)std::strlen(s);
      for (int i = 0; i < 16; ++i) tri[c][i] = -1;
      for (int i = 0; i < n; ++i) {
        char ch = s[i];
        tri[c][i] = (int8_t)(ch <= '9' ? ch - '0' : ch - 'a' + 10);
      }
      len[c] = (uint8_t)(n / 3);
      ntri[c] = len[c];
    }
    ntri[91] = 4;
  }
};
const McTables& tables() {
  static McTables t;
  return t;
}

struct Cube {
  uint32_t f;  // fine index fx + N*fy + N^2*fz (GMorpho.cu:6227-6234)
  uint8_t v;
};

// Compacted cube list in the reference's order: Morton-sorted MCM cells, 8
// cubes per cell in sub-voxel order, mixed cubes only (GMorpho.cu:6837-6878).
std::vector<Cube> mc_cubes(const uint8_t* C, int R) {
  std::vector<uint32_t> M = contour_list(C, R, 2);
  int N = 2 * R;
  std::vector<Cube> out;
  out.reserve(M.size() * 3);
  for (uint32_t c : M) {
    int x = c % R, y = (c / R) % R, z = c / (R * R);
    for (int j = 0; j < 8; ++j) {
      int fx = 2 * x + offx(j), fy = 2 * y + offy(j), fz = 2 * z + offz(j);
      int v = mc_case(C, R, fx, fy, fz);
      if (mixed(v)) out.push_back({(uint32_t)fx + (uint32_t)N * ((uint32_t)fy + (uint32_t)N * fz), (uint8_t)v});
    }
  }
  return out;
}

inline int nverts(uint16_t e) { return ((e & 1) ? 1 : 0) + ((e & 8) ? 1 : 0) + ((e & 256) ? 1 : 0); }

// findIntersection (MarchingCubesMesher.cu:98-100) on {0,1} corner values:
// p1 + (iso - v1) * (p2 - p1) / (v2 - v1), evaluated in float without
// contraction.
inline float lerp_edge(float p1, float p2, float v1, float v2, float iso) {
  volatile float d = p2 - p1;
  volatile float t = iso - v1;
  volatile float m = t * d;
  volatile float q = m / (v2 - v1);
  return p1 + q;
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }

void orc_grid_attributes(const float* P, int nv, int base_res, float margin, float* bmin3,
                         float* bmax3, float* cell, int* data_res3) {
  // Voxelizer.cu:431-455 (float arithmetic, as the reference computes it).
  float mn[3] = {P[0], P[1], P[2]}, mx[3] = {P[0], P[1], P[2]};
  for (int i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      float v = P[3 * i + a];
      if (v > mx[a]) mx[a] = v;
      if (v < mn[a]) mn[a] = v;
    }
  float thick = std::max(std::fabs(mx[0] - mn[0]), std::max(std::fabs(mx[1] - mn[1]), std::fabs(mx[2] - mn[2]))) /
                (float)base_res;
  float off = 2 * thick + margin;
  for (int a = 0; a < 3; ++a) { mn[a] -= off; mx[a] += off; }
  float cs = std::max(std::fabs(mx[0] - mn[0]), std::max(std::fabs(mx[1] - mn[1]), std::fabs(mx[2] - mn[2]))) /
             (float)base_res;
  for (int a = 0; a < 3; ++a) {
    bmin3[a] = mn[a];
    bmax3[a] = mx[a];
    data_res3[a] = (int)std::ceil((mx[a] - mn[a]) / cs);
  }
  *cell = cs;
}

void orc_voxelize(const float* P, int nv, const uint32_t* T, int nt, const float* bmin3,
                  float maxext, int base_res, uint8_t* cells) {
  const int N = 2 * base_res;
  const i64 S = 256;  // sub-voxel fixed point
  // Vertex snap: v = (P - bmin) * (N / maxext) in double, rounded half up to
  // 1/256 voxel. (The reference's rescale of Voxelizer.cu:596 and vol_vox.vert
  // map the same box onto [0,N) in float.)
  const double sc = (double)N / (double)maxext;
  std::vector<i64> X((size_t)nv * 3);
  for (int i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      double v = ((double)P[3 * i + a] - (double)bmin3[a]) * sc;
      X[3 * i + a] = (i64)std::floor(v * 256.0 + 0.5);
    }
  const size_t W = (size_t)N / 64 + 1;  // u64 words per x row
  std::vector<u64> mark((size_t)N * N * W, 0), shell((size_t)N * N * W, 0);
  auto bit = [&](std::vector<u64>& g, int x, int y, int z) -> void {
    g[((size_t)z * N + y) * W + (x >> 6)] ^= 1ull << (x & 63);
  };
  auto setb = [&](std::vector<u64>& g, int x, int y, int z) -> void {
    g[((size_t)z * N + y) * W + (x >> 6)] |= 1ull << (x & 63);
  };
  auto fdiv = [](i64 a, i64 b) -> i64 {  // floor(a/b), b > 0
    i64 q = a / b;
    if ((a % b != 0) && (a < 0)) --q;
    return q;
  };
  for (int t = 0; t < nt; ++t) {
    const i64* A = &X[3 * (size_t)T[3 * t]];
    const i64* B = &X[3 * (size_t)T[3 * t + 1]];
    const i64* C = &X[3 * (size_t)T[3 * t + 2]];
    // ---- volume pass (vol_vox_bit.frag:35-52): orthographic along z, each
    // covered pixel centre toggles the column below the triangle's depth.
    {
      const i64 *a = A, *b = B, *c = C;
      i64 area = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
      if (area != 0) {
        if (area < 0) { std::swap(b, c); area = -area; }
        i64 xmn = std::min(a[0], std::min(b[0], c[0])), xmx = std::max(a[0], std::max(b[0], c[0]));
        i64 ymn = std::min(a[1], std::min(b[1], c[1])), ymx = std::max(a[1], std::max(b[1], c[1]));
        i64 i0 = std::max<i64>(0, fdiv(xmn - S / 2 + S - 1, S)), i1 = std::min<i64>(N - 1, fdiv(xmx - S / 2, S));
        i64 j0 = std::max<i64>(0, fdiv(ymn - S / 2 + S - 1, S)), j1 = std::min<i64>(N - 1, fdiv(ymx - S / 2, S));
        const i64* V[3] = {a, b, c};
        for (i64 j = j0; j <= j1; ++j)
          for (i64 i = i0; i <= i1; ++i) {
            i64 px = i * S + S / 2, py = j * S + S / 2;
            i64 w[3];
            bool in = true;
            for (int e = 0; e < 3 && in; ++e) {
              const i64* p0 = V[e];
              const i64* p1 = V[(e + 1) % 3];
              i64 ex = p1[0] - p0[0], ey = p1[1] - p0[1];
              i64 f = ex * (py - p0[1]) - ey * (px - p0[0]);
              bool tl = (ey < 0) || (ey == 0 && ex < 0);
              if (f < 0 || (f == 0 && !tl)) in = false;
              w[(e + 2) % 3] = f;  // weight of the vertex opposite edge e
            }
            if (!in) continue;
            i64 num = w[0] * a[2] + w[1] * b[2] + w[2] * c[2];
            i64 zc = fdiv(num, area * S);
            if (zc < 0) continue;
            if (zc > N - 1) zc = N - 1;
            bit(mark, (int)i, (int)j, (int)zc);
          }
      }
    }
    // ---- surface pass (surf_vox_conserv.geo:41-122, *_bit.frag:35-47).
    {
      i64 e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
      i64 e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
      i64 n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      typedef unsigned __int128 u128;
      u128 a2 = (u128)((__int128)n[0] * n[0]) + (u128)((__int128)n[1] * n[1]) + (u128)((__int128)n[2] * n[2]);
      if (a2 == 0) continue;
      i64 d0[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
      i64 d1[3] = {B[0] - C[0], B[1] - C[1], B[2] - C[2]};
      i64 l0 = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2];
      i64 l1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
      i64 l2 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
      u128 ml = (u128)std::max(l0, std::max(l1, l2));
      // emit iff maxl2^2 / |n|^2 < 50^2 (Voxelizer.cu:680)
      if (!(ml * ml < (u128)2500 * a2)) continue;
      i64 an[3] = {std::llabs(n[0]), std::llabs(n[1]), std::llabs(n[2])};
      i64 mxn = std::max(an[0], std::max(an[1], an[2]));
      int ax = an[0] == mxn ? 0 : (an[1] == mxn ? 1 : 2);
      // dominant plane axes (u,v) and depth axis d (geo: .yzx, .zxy, .xyz)
      int ua = ax == 0 ? 1 : (ax == 1 ? 2 : 0);
      int va = ax == 0 ? 2 : (ax == 1 ? 0 : 1);
      const i64* V[3] = {A, B, C};
      i64 area = (B[ua] - A[ua]) * (C[va] - A[va]) - (B[va] - A[va]) * (C[ua] - A[ua]);
      if (area < 0) std::swap(V[1], V[2]);
      i64 umn = std::min(A[ua], std::min(B[ua], C[ua])), umx = std::max(A[ua], std::max(B[ua], C[ua]));
      i64 vmn = std::min(A[va], std::min(B[va], C[va])), vmx = std::max(A[va], std::max(B[va], C[va]));
      i64 iu0 = std::max<i64>(0, fdiv(umn - S + S - 1, S)), iu1 = std::min<i64>(N - 1, fdiv(umx, S));
      i64 iv0 = std::max<i64>(0, fdiv(vmn - S + S - 1, S)), iv1 = std::min<i64>(N - 1, fdiv(vmx, S));
      i64 nd = n[ax];
      i64 den = nd * S;
      for (i64 iv = iv0; iv <= iv1; ++iv)
        for (i64 iu = iu0; iu <= iu1; ++iu) {
          i64 ulo = iu * S, uhi = ulo + S, vlo = iv * S, vhi = vlo + S;
          bool ok = true;
          for (int e = 0; e < 3 && ok; ++e) {
            const i64* p0 = V[e];
            const i64* p1 = V[(e + 1) % 3];
            i64 eu = p1[ua] - p0[ua], ev = p1[va] - p0[va];
            i64 pu = (-ev > 0) ? uhi : ulo;
            i64 pv = (eu > 0) ? vhi : vlo;
            i64 f = eu * (pv - p0[va]) - ev * (pu - p0[ua]);
            if (f < 0) ok = false;
          }
          if (!ok) continue;
          i64 pu = ulo + S / 2, pv = vlo + S / 2;
          i64 num = A[ax] * nd - (n[ua] * (pu - A[ua]) + n[va] * (pv - A[va]));
          i64 q = den > 0 ? fdiv(num, den) : fdiv(-num, -den);
          if (q < 0 || q >= N) continue;
          int c3[3];
          c3[ua] = (int)iu; c3[va] = (int)iv; c3[ax] = (int)q;
          setb(shell, c3[0], c3[1], c3[2]);
        }
    }
  }
  // Parity fill: voxel k of a column is set when an odd number of markers lie
  // at or above it; then OR the shell and pack into cells (Voxelizer.cu:798-849).
  const int R = base_res;
  std::memset(cells, 0, (size_t)R * R * R);
  for (int y = 0; y < N; ++y)
    for (size_t w = 0; w < W; ++w) {
      u64 acc = 0;
      for (int z = N - 1; z >= 0; --z) {
        size_t idx = ((size_t)z * N + y) * W + w;
        acc ^= mark[idx];
        mark[idx] = acc | shell[idx];
      }
    }
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x)
        if ((mark[((size_t)z * N + y) * W + (x >> 6)] >> (x & 63)) & 1)
          cells[cidx(R, x >> 1, y >> 1, z >> 1)] |= (uint8_t)(1 << ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)));
}

void orc_voxelize_multislice(const float* P, int nv, const uint32_t* T, int nt, const float* bmin3, float maxext,
                             int base_res, uint8_t* cells) {
  // Voxelizer::VoxelizeByMultiSlicing (Voxelizer.cu:375-400, kernel
  // MultiSlicing at 188-331): rays along +x through the centres of the fine
  // voxel columns (fine y, z) of each cell row (Voxelizer.cu:262-266); a
  // triangle covering a ray's centre under the top-left rule (the FLT_MIN
  // bias of the edge tests, 252-255) toggles every voxel of the column from
  // the one holding the crossing on (qi <= x, 288-306). No surface shell.
  // Restated on the snapped vertices of orc_voxelize with exact integer edge
  // functions and depths; a crossing before the grid toggles the whole
  // column, one past it nothing.
  const int N = 2 * base_res;
  const i64 S = 256;
  const double sc = (double)N / (double)maxext;
  std::vector<i64> X((size_t)nv * 3);
  for (int i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      double v = ((double)P[3 * i + a] - (double)bmin3[a]) * sc;
      X[3 * i + a] = (i64)std::floor(v * 256.0 + 0.5);
    }
  auto fdiv = [](i64 a, i64 b) -> i64 {  // floor(a/b), b > 0
    i64 q = a / b;
    if ((a % b != 0) && (a < 0)) --q;
    return q;
  };
  // one toggle bit per (x, y, z): the column toggles from x on
  std::vector<uint8_t> mark((size_t)N * N * N, 0);
  for (int t = 0; t < nt; ++t) {
    const i64* a = &X[3 * (size_t)T[3 * t]];
    const i64* b = &X[3 * (size_t)T[3 * t + 1]];
    const i64* c = &X[3 * (size_t)T[3 * t + 2]];
    // projection on the (y, z) plane, counter-clockwise
    i64 area = (b[1] - a[1]) * (c[2] - a[2]) - (b[2] - a[2]) * (c[1] - a[1]);
    if (area == 0) continue;
    if (area < 0) {
      std::swap(b, c);
      area = -area;
    }
    const i64* V[3] = {a, b, c};
    i64 ymn = std::min(a[1], std::min(b[1], c[1])), ymx = std::max(a[1], std::max(b[1], c[1]));
    i64 zmn = std::min(a[2], std::min(b[2], c[2])), zmx = std::max(a[2], std::max(b[2], c[2]));
    i64 j0 = std::max<i64>(0, fdiv(ymn - S / 2 + S - 1, S)), j1 = std::min<i64>(N - 1, fdiv(ymx - S / 2, S));
    i64 k0 = std::max<i64>(0, fdiv(zmn - S / 2 + S - 1, S)), k1 = std::min<i64>(N - 1, fdiv(zmx - S / 2, S));
    for (i64 k = k0; k <= k1; ++k)
      for (i64 j = j0; j <= j1; ++j) {
        const i64 py = j * S + S / 2, pz = k * S + S / 2;
        i64 w[3];
        bool in = true;
        for (int e = 0; e < 3 && in; ++e) {
          const i64* p0 = V[e];
          const i64* p1 = V[(e + 1) % 3];
          const i64 ey = p1[1] - p0[1], ez = p1[2] - p0[2];
          const i64 f = ey * (pz - p0[2]) - ez * (py - p0[1]);
          const bool tl = (ez < 0) || (ez == 0 && ey < 0);
          if (f < 0 || (f == 0 && !tl)) in = false;
          w[(e + 2) % 3] = f;
        }
        if (!in) continue;
        // the voxel holding the crossing: floor(x / S) of the plane's x
        i64 xq = fdiv(w[0] * a[0] + w[1] * b[0] + w[2] * c[0], area * S);
        if (xq >= N) continue;
        if (xq < 0) xq = 0;
        mark[((size_t)k * N + (size_t)j) * N + (size_t)xq] ^= 1;
      }
  }
  const int R = base_res;
  std::memset(cells, 0, (size_t)R * R * R);
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y) {
      uint8_t acc = 0;
      for (int x = 0; x < N; ++x) {
        acc ^= mark[((size_t)z * N + y) * N + x];
        if (acc) cells[cidx(R, x >> 1, y >> 1, z >> 1)] |= (uint8_t)(1 << ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)));
      }
    }
}

void orc_scale_from_lattice(const float* lattice, int L, int N, int8_t* out) {
  // Fixed-point trilinear sampling: lattice node i sits on fine voxel i*N/L
  // (N/L a power of two >= 1), values are first rounded to 1/256, samples use
  // clamp-to-edge, and the radius is floor() of the interpolant.
  const int s = N / L;
  std::vector<i64> q((size_t)L * L * L);
  for (size_t i = 0; i < q.size(); ++i) q[i] = (i64)std::floor((double)lattice[i] * 256.0 + 0.5);
  auto Q = [&](int x, int y, int z) {
    x = std::min(x, L - 1); y = std::min(y, L - 1); z = std::min(z, L - 1);
    return q[(size_t)x + (size_t)L * ((size_t)y + (size_t)L * z)];
  };
  const i64 den = (i64)s * s * s * 256;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        int ix = x / s, iy = y / s, iz = z / s;
        i64 fx = x - ix * s, fy = y - iy * s, fz = z - iz * s;
        i64 acc = 0;
        for (int k = 0; k < 8; ++k) {
          i64 wx = (k & 1) ? fx : s - fx, wy = (k & 2) ? fy : s - fy, wz = (k & 4) ? fz : s - fz;
          acc += wx * wy * wz * Q(ix + (k & 1), iy + ((k >> 1) & 1), iz + ((k >> 2) & 1));
        }
        i64 v = acc >= 0 ? acc / den : -((-acc + den - 1) / den);
        v = std::max<i64>(-128, std::min<i64>(127, v));
        out[(size_t)x + (size_t)N * ((size_t)y + (size_t)N * z)] = (int8_t)v;
      }
}

void orc_pyramid(const uint8_t* cells, int R, const int8_t* scale, int r, uint64_t* out) {
  Radius rad{scale, r, 2 * R};
  Pyramid py = build_pyramid(cells, R, rad);
  size_t o = 0;
  for (int L = 0; L < py.levels; ++L) {
    std::memcpy(out + o, py.lv[L].data(), py.lv[L].size() * sizeof(u64));
    o += py.lv[L].size();
  }
}

void orc_dilate(const uint8_t* I, int R, const int* data_res3, const int8_t* scale, int r, int se,
                int brute, uint8_t* D) {
  Radius rad{scale, r, 2 * R};
  std::memcpy(D, I, (size_t)R * R * R);
  Pyramid py;
  if (!brute) py = build_pyramid(I, R, rad);
  dilate_cells(I, R, data_res3, rad, se, brute, &py, D, 0, R, true);
}

void orc_contour_K(const uint8_t* D, int R, uint8_t* K) { contour_K(D, R, K); }

void orc_erode(const uint8_t* I, const uint8_t* D, int R, const int* data_res3,
               const int8_t* scale, int r, int brute, uint8_t* C) {
  Radius rad{scale, r, 2 * R};
  size_t n = (size_t)R * R * R;
  std::vector<uint8_t> K(n);
  contour_K(D, R, K.data());
  Pyramid pk;
  if (!brute) pk = build_pyramid(K.data(), R, rad);
  std::memcpy(C, D, n);
  erode_cells(I, D, K.data(), R, data_res3, rad, brute, &pk, C, 0, R, true);
}

int64_t orc_contour_list(const uint8_t* G, int R, int which, uint32_t* out) {
  std::vector<uint32_t> l = contour_list(G, R, which);
  if (out) std::memcpy(out, l.data(), l.size() * sizeof(uint32_t));
  return (int64_t)l.size();
}

void orc_mc_tables(uint16_t* edge256, uint8_t* ntri256, int8_t* tri4096) {
  const McTables& t = tables();
  std::memcpy(edge256, t.edge, sizeof(t.edge));
  std::memcpy(ntri256, t.ntri, sizeof(t.ntri));
  std::memcpy(tri4096, t.tri, sizeof(t.tri));
}

void orc_mc_counts(const uint8_t* C, int R, uint32_t* nv, uint32_t* nt) {
  const McTables& tb = tables();
  std::vector<Cube> cubes = mc_cubes(C, R);
  uint64_t v = 0, t = 0;
  for (auto& c : cubes) {
    v += nverts(tb.edge[c.v]);
    t += tb.ntri[c.v];
  }
  *nv = (uint32_t)v;
  *nt = (uint32_t)t;
}

void orc_mc(const uint8_t* C, int R, const float* bmin3, float cell, float* V, float* Nrm,
            uint32_t* T) {
  const McTables& tb = tables();
  std::vector<Cube> cubes = mc_cubes(C, R);
  const int N = 2 * R;
  size_t n = cubes.size();
  // exclusive scans of per-cube vertex / triangle counts (MCM.cu:1544-1571)
  std::vector<uint32_t> sv(n), st(n);
  uint32_t av = 0, at = 0;
  for (size_t i = 0; i < n; ++i) {
    sv[i] = av;
    st[i] = at;
    av += nverts(tb.edge[cubes[i].v]);
    at += tb.ntri[cubes[i].v];
  }
  // fine cube index -> compact id (SplatIndices + BuildNeighbors, GMorpho.cu:6376, 6459-6496)
  std::vector<std::pair<uint32_t, uint32_t>> idmap(n);
  for (size_t i = 0; i < n; ++i) idmap[i] = {cubes[i].f, (uint32_t)i};
  std::sort(idmap.begin(), idmap.end());
  auto lookup = [&](uint32_t f) -> uint32_t {
    auto it = std::lower_bound(idmap.begin(), idmap.end(), std::make_pair(f, 0u));
    if (it != idmap.end() && it->first == f) return it->second;
    return 0xffffffffu;
  };
  const float h = 0.5f * cell;
  const float iso = 0.95f;
  std::vector<double> acc((size_t)av * 3, 0.0);
  struct Late {
    uint32_t slot;
    uint32_t t[3];
  };
  std::vector<Late> late;
  uint32_t extra[3 * 16];
  for (size_t i = 0; i < n; ++i) {
    uint32_t f = cubes[i].f;
    uint8_t cv = cubes[i].v;
    int fx = f % N, fy = (f / N) % N, fz = f / ((uint32_t)N * N);
    uint32_t id[8];
    // neighbour cubes at (1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)
    const int no[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int k = 0; k < 8; ++k)
      id[k] = lookup((uint32_t)(fx + no[k][0]) + (uint32_t)N * ((uint32_t)(fy + no[k][1]) + (uint32_t)N * (fz + no[k][2])));
    float px = std::fma((float)fx, h, bmin3[0]);
    float py = std::fma((float)fy, h, bmin3[1]);
    float pz = std::fma((float)fz, h, bmin3[2]);
    float v[8];
    for (int k = 0; k < 8; ++k) v[k] = ((cv >> k) & 1) ? 1.f : 0.f;
    uint16_t e = tb.edge[cv];
    uint32_t vert[12] = {0};
    uint32_t ind = 0;
    // createTrianglesAndVertices3D (MarchingCubesMesher.cu:1312-1500)
    if (e & 1) {
      vert[0] = sv[i] + ind++;
      float* o = &V[3 * (size_t)vert[0]];
      o[0] = lerp_edge(px, px + h, v[0], v[1], iso); o[1] = py; o[2] = pz;
    }
    if ((e & 2) && id[1] != 0xffffffffu) vert[1] = sv[id[1]] + ((tb.edge[cubes[id[1]].v] & 1) ? 1 : 0);
    if (e & 4) vert[2] = id[3] != 0xffffffffu ? sv[id[3]] : 0;
    if (e & 8) {
      vert[3] = sv[i] + ind++;
      float* o = &V[3 * (size_t)vert[3]];
      o[0] = px; o[1] = lerp_edge(py + h, py, v[3], v[0], iso); o[2] = pz;
    }
    if (e & 16) vert[4] = id[4] != 0xffffffffu ? sv[id[4]] : 0;
    if ((e & 32) && id[5] != 0xffffffffu) vert[5] = sv[id[5]] + ((tb.edge[cubes[id[5]].v] & 1) ? 1 : 0);
    if (e & 64) vert[6] = id[7] != 0xffffffffu ? sv[id[7]] : 0;
    if ((e & 128) && id[4] != 0xffffffffu) vert[7] = sv[id[4]] + ((tb.edge[cubes[id[4]].v] & 1) ? 1 : 0);
    if (e & 256) {
      vert[8] = sv[i] + ind++;
      float* o = &V[3 * (size_t)vert[8]];
      o[0] = px; o[1] = py; o[2] = lerp_edge(pz, pz + h, v[0], v[4], iso);
    }
    const int nb[3] = {1, 2, 3};
    for (int q = 0; q < 3; ++q) {
      int ebit = 512 << q;
      if (!(e & ebit)) continue;
      uint32_t nid = id[nb[q]];
      uint16_t e1 = nid != 0xffffffffu ? tb.edge[cubes[nid].v] : 0;
      uint32_t d = ((e1 & 1) ? 1 : 0) + ((e1 & 8) ? 1 : 0);
      vert[9 + q] = nid != 0xffffffffu ? sv[nid] + d : d;
    }
    for (int q = 0; q < tb.len[cv]; ++q) {
      uint32_t* o = q < tb.ntri[cv] ? &T[3 * ((size_t)st[i] + q)] : &extra[3 * (size_t)(q - tb.ntri[cv])];
      o[0] = vert[tb.tri[cv][3 * q + 2]];
      o[1] = vert[tb.tri[cv][3 * q + 1]];
      o[2] = vert[tb.tri[cv][3 * q]];
    }
    for (int q = tb.ntri[cv]; q < tb.len[cv]; ++q)
      late.push_back({st[i] + (uint32_t)q, {extra[3 * (q - tb.ntri[cv])], extra[3 * (q - tb.ntri[cv]) + 1],
                                            extra[3 * (q - tb.ntri[cv]) + 2]}});
  }
  // Triangles past a cube's counted slots overwrite the following slots.
  for (auto& l : late)
    if (l.slot < at) std::memcpy(&T[3 * (size_t)l.slot], l.t, sizeof(l.t));
  // Area-weighted vertex normals (GMorpho.cu:6498-6537), accumulated in double.
  for (uint32_t t = 0; t < at; ++t) {
    const uint32_t* tr = &T[3 * (size_t)t];
    const float* p0 = &V[3 * (size_t)tr[0]];
    const float* p1 = &V[3 * (size_t)tr[1]];
    const float* p2 = &V[3 * (size_t)tr[2]];
    double a[3] = {(double)p1[0] - p0[0], (double)p1[1] - p0[1], (double)p1[2] - p0[2]};
    double b[3] = {(double)p2[0] - p0[0], (double)p2[1] - p0[1], (double)p2[2] - p0[2]};
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 3; ++q) acc[3 * (size_t)tr[k] + q] += c[q];
  }
  for (uint32_t i = 0; i < av; ++i) {
    double* a = &acc[3 * (size_t)i];
    double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int q = 0; q < 3; ++q) Nrm[3 * (size_t)i + q] = (float)(l > 0 ? a[q] / l : 0.0);
  }
}


// BilateralMeshFiltering + BilateralMeshFilteringNormalize (GMorpho.cu:6539-6620)
// iterated as ExtractClosingMesh does (GMorpho.cu:6957-6978): sigma_p = 1 and
// sigma_n = 1e10 for the first two iterations, 0.2 after. Per triangle and
// corner v with other corners a, b (in corner order), sumN = w(v,a) n_a +
// w(v,b) n_b with w(v,u) = exp(-(|p_v-p_u|^2 / sigma_p^2 + |n_v-n_u|^2 /
// sigma_n^2)) (exp as exp_weight); the sums of a vertex are added in triangle
// order (the reference's float atomics, serialized), then normalized. A zero
// sum stays zero (the reference would divide by zero).
// exp(x) of a bilateral weight (x <= 0) in float: Cody-Waite reduction
// x = k ln2 + r with ln2 split as 355/512 + (ln2 - 355/512) (k ln2_hi exact
// for |k| < 2^15), exp(r) by its Taylor series to degree 7 in Horner form
// (|r| <= 0.35: truncation below 1e-8 relative), times 2^k. Below -87 the
// result would leave the normal range and is 0. The HIP kernel evaluates the
// same operations in the same order, so both give the same bits; the test
// suite holds this function to numpy's exp within a few ulp.
static float exp_weight(float x) {
  if (!(x >= -87.f)) return 0.f;
  if (x > 0.f) x = 0.f;
  const float k = std::nearbyint((float)(x * 1.44269504f));
  volatile float a = k * 0.693359375f;
  volatile float b = k * -2.12194440e-4f;
  volatile float r0 = x - (float)a;
  volatile float r = (float)r0 - (float)b;
  const float c[8] = {1.f / 5040.f, 1.f / 720.f, 1.f / 120.f, 1.f / 24.f, 1.f / 6.f, 0.5f, 1.f, 1.f};
  volatile float p = c[0];
  for (int i = 1; i < 8; ++i) {
    volatile float m = (float)p * (float)r;
    p = (float)m + c[i];
  }
  return std::ldexp((float)p, (int)k);
}

void orc_exp_weight(const float* x, float* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = exp_weight(x[i]);
}

void orc_bilateral_normals(const float* V, float* Nrm, const uint32_t* T, int nv, int nt, int iters) {
  std::vector<float> acc((size_t)3 * nv);
  auto d2 = [](const float* a, const float* b) {
    volatile float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    volatile float xx = x * x, yy = y * y, zz = z * z;
    volatile float s = xx + yy;
    return (float)(s + zz);
  };
  for (int it = 0; it < iters; ++it) {
    const float sp = 1.f, sn = it < 2 ? 1e10f : 0.2f;
    const float sp2 = sp * sp, sn2 = sn * sn;
    std::fill(acc.begin(), acc.end(), 0.f);
    for (int t = 0; t < nt; ++t) {
      const uint32_t* tri = T + 3 * (size_t)t;
      for (int c = 0; c < 3; ++c) {
        const uint32_t v = tri[c], a = tri[c == 0 ? 1 : 0], b = tri[c == 2 ? 1 : 2];
        const float *pv = V + 3 * (size_t)v, *pa = V + 3 * (size_t)a, *pb = V + 3 * (size_t)b;
        const float *nv_ = Nrm + 3 * (size_t)v, *na = Nrm + 3 * (size_t)a, *nb = Nrm + 3 * (size_t)b;
        volatile float ea = d2(pv, pa) / sp2 + d2(nv_, na) / sn2;
        volatile float eb = d2(pv, pb) / sp2 + d2(nv_, nb) / sn2;
        const float wa = exp_weight(-(float)ea), wb = exp_weight(-(float)eb);
        for (int q = 0; q < 3; ++q) {
          volatile float s = wa * na[q];
          volatile float u = wb * nb[q];
          volatile float sum = s + u;
          acc[3 * (size_t)v + q] += sum;
        }
      }
    }
    for (int i = 0; i < nv; ++i) {
      float* a = &acc[3 * (size_t)i];
      volatile float xx = a[0] * a[0], yy = a[1] * a[1], zz = a[2] * a[2];
      volatile float s = xx + yy;
      const float l = std::sqrt((float)(s + zz));
      for (int q = 0; q < 3; ++q) Nrm[3 * (size_t)i + q] = l > 0.f ? a[q] / l : a[q];
    }
  }
}


// ScaleField::UpdateScaleGrid (ScaleField.cu:81-141) for one brush point over
// every fine voxel, with tukey (cuda_math.h:696-704) and the edit modes of
// ScaleField.h:32 (1 MAX, 2 MIN, 3 ZERO). The voxel's sample point is its
// corner bbox_min + cell * id; h = (support - scale) / sqrt(1 - sqrt(0.5)),
// all in float. Voxels at distance >= scale + h keep their value.
void orc_scale_brush(int8_t* grid, int N, const float* bmin3, float cell, const float* pos3, float scale,
                     float support, float global_scale, int mode) {
  const float v = scale;
  volatile float rt = std::sqrt(1.f - std::sqrt(0.5f));
  volatile float h = (support - v) / (float)rt;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        const int id[3] = {x, y, z};
        float d2 = 0.f;
        for (int a = 0; a < 3; ++a) {
          volatile float m = cell * (float)id[a];
          volatile float u = m + bmin3[a];
          volatile float r = u - pos3[a];
          volatile float rr = r * r;
          volatile float acc = d2 + rr;
          d2 = acc;
        }
        const float dist = std::sqrt(d2);
        volatile float reach = v + (float)h;
        if (!(dist < reach)) continue;
        float distrib;
        if (dist < v) {
          distrib = 1.f;
        } else {
          volatile float q = (dist - v) / (float)h;
          volatile float q2 = q * q;
          volatile float o = 1.f - q2;
          distrib = o * o;
        }
        int8_t& g = grid[(size_t)x + (size_t)N * ((size_t)y + (size_t)N * z)];
        const float prev = (float)g;
        const float target = mode == 3 ? global_scale / cell : scale / cell;
        volatile float keep = prev * (1.f - distrib);
        volatile float add = distrib * target;
        float d = std::floor((float)(keep + add));
        d = std::max(std::min(d, 127.f), 0.f);
        if (mode == 1) g = (int8_t)std::floor(std::max(d, prev));
        else if (mode == 2) g = (int8_t)std::floor(std::min(d, prev));
        else if (mode == 3) g = (int8_t)d;
      }
}

int64_t orc_stage_slab(int stage, const uint8_t* I, const uint8_t* D, const uint8_t* C, int R,
                       const int* data_res3, const int8_t* scale, int r, int z0, int z1,
                       uint8_t* out) {
  Radius rad{scale, r, 2 * R};
  size_t n = (size_t)R * R * R;
  if (stage == 0) {
    std::memcpy(out, I, n);
    Pyramid py = build_pyramid(I, R, rad);
    dilate_cells(I, R, data_res3, rad, 0, 0, &py, out, z0, z1, true);
    return (int64_t)(z1 - z0) * R * R;
  }
  if (stage == 1) {
    std::vector<uint8_t> K(n);
    contour_K(D, R, K.data());
    Pyramid pk = build_pyramid(K.data(), R, rad);
    std::memcpy(out, D, n);
    erode_cells(I, D, K.data(), R, data_res3, rad, 0, &pk, out, z0, z1, true);
    return (int64_t)(z1 - z0) * R * R;
  }
  if (stage == 2) {
    int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : cnt)
    for (int z = z0; z < z1; ++z)
      for (int y = 0; y < R; ++y)
        for (int x = 0; x < R; ++x)
          if (tag_mcm(C, R, x, y, z)) ++cnt;
    return cnt;
  }
  return -1;
}

}  // extern "C"
