/* GMorpho oracle: a CPU restatement of the reference's closing path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline), never as the product path.
 *
 * Grids are the reference's cell-byte layout: R^3 bytes, index x + R*(y + R*z),
 * bit k of a byte = fine voxel (2x + (k&1), 2y + ((k>>1)&1), 2z + ((k>>2)&1))
 * (Grid.h:61-68, GMorpho.cu:468-483). N = 2R fine voxels per axis.
 *
 * Radius source (a "radius spec"): scale == NULL -> uniform radius r (masked to
 * 7 bits as Grid.cu:725 does); otherwise an N^3 signed-char ScaleField grid
 * (x fastest) clamped to [2,127] per fine voxel (Grid.cu:990-992).
 *
 * brute = 0 runs the reference algorithm (depth-first descent of the
 * occupancy+radius max-mipmap); brute = 1 scans every fine voxel in the
 * structuring element's window instead. Both must agree (the DFS prunes with
 * exact bounds).
 */
#ifndef GMORPHO_ORACLE_H
#define GMORPHO_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Voxelizer::ComputeGridAttributes (Voxelizer.cu:431-455) with the margin of
 * GMorpho::Load (GMorpho.cu:7737) supplied by the caller. */
void orc_grid_attributes(const float* P, int nv, int base_res, float margin, float* bmin3,
                         float* bmax3, float* cell, int* data_res3);

/* Voxelizer::VoxelizeConservative (Voxelizer.cu:724-796) restated with the
 * fixed-point rule: parity (XOR) solid fill along z at pixel centres plus a
 * conservative surface shell on each triangle's dominant plane. */
void orc_voxelize(const float* P, int nv, const uint32_t* T, int nt, const float* bmin3,
                  float maxext, int base_res, uint8_t* cells);

/* Voxelizer::VoxelizeByMultiSlicing (Voxelizer.cu:188-331, 375-400; not
 * called by the reference's GMorpho::Load, GMorpho.cu:7741): parity along x
 * from each covered fine column's crossing voxel on, on the same snapped
 * vertices and exact integer rule as orc_voxelize; no surface shell. */
void orc_voxelize_multislice(const float* P, int nv, const uint32_t* T, int nt, const float* bmin3, float maxext,
                             int base_res, uint8_t* cells);

/* Config-2 style ScaleField: FP32 L^3 lattice -> N^3 signed-char radii,
 * fixed-point trilinear sampling then floor, clamped to [-128,127]. */
void orc_scale_from_lattice(const float* lattice, int L, int N, int8_t* out);

/* Grid::BuildTexMipmaps (Grid.cu:643-1154): the 64-bit occupancy + 7-bit
 * radius max-mipmap, levels 0..log2(R) concatenated, level L is
 * (R>>L)^3 texels x fastest. R must be a power of two. */
void orc_pyramid(const uint8_t* cells, int R, const int8_t* scale, int r, uint64_t* out);

/* Two-pass sphere (se=0) or cube (se=1) dilation (GMorpho.cu:2828-3056,
 * 1838-1934, 2603-2721, driver 3538-3596). D must hold a copy of I on entry
 * semantics are applied internally; I and D may not alias. */
void orc_dilate(const uint8_t* I, int R, const int* data_res3, const int8_t* scale, int r, int se,
                int brute, uint8_t* D);

/* Grid::TransformToContour (Grid.cu:2381-2408, kernel 1230-1327). */
void orc_contour_K(const uint8_t* D, int R, uint8_t* K);

/* Two-pass ball erosion (GMorpho.cu:4349-4648, 5041-5374, driver 6627-6702). */
void orc_erode(const uint8_t* I, const uint8_t* D, int R, const int* data_res3,
               const int8_t* scale, int r, int brute, uint8_t* C);

/* Contour cell lists, Morton-sorted, as the reference builds them:
 * which = 0 tight (Grid.cu:1572), 1 base 6-connected (1499), 2 MCM (1329).
 * Returns the count; writes cell indices when out != NULL. */
int64_t orc_contour_list(const uint8_t* G, int R, int which, uint32_t* out);

/* ExtractClosingMesh (GMorpho.cu:6814-7001) + MarchingCubesMesher::createMesh3D
 * (MarchingCubesMesher.cu:1297-1618): counts, then vertices / area-weighted
 * normals / triangles in the reference's output order. */
void orc_mc_counts(const uint8_t* C, int R, uint32_t* nv, uint32_t* nt);
void orc_mc(const uint8_t* C, int R, const float* bmin3, float cell, float* V, float* Nrm,
            uint32_t* T);

/* The MC case tables the oracle uses (edge mask per case, triangle count per
 * case, 16 edge ids per case terminated by -1). */
void orc_mc_tables(uint16_t* edge256, uint8_t* ntri256, int8_t* tri4096);

/* Bilateral normal filtering of ExtractClosingMesh (GMorpho.cu:6539-6620,
 * 6957-6978): iters iterations on the unit normals Nrm (in place), each
 * vertex's sums added in triangle order. */
void orc_bilateral_normals(const float* V, float* Nrm, const uint32_t* T, int nv, int nt, int iters);
/* The float exp of the bilateral weights (x <= 0; 0 below -87). */
void orc_exp_weight(const float* x, float* out, int n);

/* ScaleField::UpdateScaleGrid (ScaleField.cu:81-141): one brush point applied
 * to the N^3 signed-char radius grid (x fastest) whose voxel i has its corner
 * at bmin + cell * i; mode 1 MAX_BRUSH, 2 MIN_BRUSH, 3 ZERO_BRUSH
 * (ScaleField.h:32). */
void orc_scale_brush(int8_t* grid, int N, const float* bmin3, float cell, const float* pos3, float scale,
                     float support, float global_scale, int mode);

/* Timed sub-steps for the CPU baseline: run the named stage on the cell slab
 * z in [z0,z1) only. stage 0 = coarse+fine dilation of those cells,
 * 1 = coarse+fine erosion of those cells, 2 = MC classification/counting.
 * Returns the number of cells processed. */
int64_t orc_stage_slab(int stage, const uint8_t* I, const uint8_t* D, const uint8_t* C, int R,
                       const int* data_res3, const int8_t* scale, int r, int z0, int z1,
                       uint8_t* out);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
