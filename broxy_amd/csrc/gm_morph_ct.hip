// Streaming morphology kernels for one uniform radius each: the chord table
// of the radius is compiled into the kernel (LDS reads at immediate offsets,
// unrolled classes), so each kernel carries only its own radius's code and
// register budget. This file is compiled kCtParts times with
// -DGM_CT_PART=p; part p holds the radii rho with rho % kCtParts == p and
// registers them with ct_register_part<p>.
#include "gm_morph_dev.h"

#ifndef GM_CT_PART
#error "compile with -DGM_CT_PART=<part>"
#endif

namespace gm {
namespace {

template <int PASS, int RHO>
__global__ void __launch_bounds__(64 * kMaxWaves) k_pass_ct(MorphArgs a) {
  run_pass<PASS, RHO>(a);
}

// Many jobs of radius RHO in one launch (entries name job and tile).
template <int PASS, int RHO>
__global__ void __launch_bounds__(64 * kMaxWaves) k_pass_ct_b(const MorphArgs* __restrict__ tab,
                                                               const u32* __restrict__ queue, int qcOff, int qlOff,
                                                               int qCap) {
  run_pass_batched<PASS, RHO>(tab, queue, qcOff, qlOff, qCap);
}

// Dilation uses rho = r + 1 <= 32 (corner tables); erosion rho = r - 1 <= 30
// (corner tables for the coarse pass, centre tables for the fine pass); the
// cube dilation rho = r <= 32 (box tables, passes 4 and 5).
template <int RHO>
void register_rho(CtKernels& k) {
  if constexpr (RHO % kCtParts == GM_CT_PART) {
    k.fn[0][RHO] = reinterpret_cast<const void*>(&k_pass_ct<0, RHO>);
    k.fn[1][RHO] = reinterpret_cast<const void*>(&k_pass_ct<1, RHO>);
    k.fnB[0][RHO] = reinterpret_cast<const void*>(&k_pass_ct_b<0, RHO>);
    k.fnB[1][RHO] = reinterpret_cast<const void*>(&k_pass_ct_b<1, RHO>);
    k.fnBox[0][RHO] = reinterpret_cast<const void*>(&k_pass_ct<4, RHO>);
    k.fnBox[1][RHO] = reinterpret_cast<const void*>(&k_pass_ct<5, RHO>);
    if constexpr (RHO <= kCtMaxRho - 2) {
      k.fn[2][RHO] = reinterpret_cast<const void*>(&k_pass_ct<2, RHO>);
      k.fn[3][RHO] = reinterpret_cast<const void*>(&k_pass_ct<3, RHO>);
      k.fnB[2][RHO] = reinterpret_cast<const void*>(&k_pass_ct_b<2, RHO>);
      k.fnB[3][RHO] = reinterpret_cast<const void*>(&k_pass_ct_b<3, RHO>);
    }
  }
}
template <int... Rs>
void register_all(CtKernels& k, std::integer_sequence<int, Rs...>) {
  (register_rho<Rs + 1>(k), ...);
}

}  // namespace

#define GM_CT_REGISTER_NAME2(p) ct_register_part##p
#define GM_CT_REGISTER_NAME(p) GM_CT_REGISTER_NAME2(p)
void GM_CT_REGISTER_NAME(GM_CT_PART)(CtKernels& k) {
  register_all(k, std::make_integer_sequence<int, kCtMaxRho>{});
}

}  // namespace gm
