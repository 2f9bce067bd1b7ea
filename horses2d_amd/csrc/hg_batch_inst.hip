// One physics/precision variant of the batched kernels for one Riemann solver: compiled once
// per variant with -DHG_BV=<2*physics + mixed> (physics 0 Euler, 1 Navier-Stokes BR1,
// 2 Navier-Stokes IP) and -DHG_RS1=<solver + 1> (0: chosen per face at run time).  The
// gradient kernels of a variant live in its HG_RS1=0 object.
#include <type_traits>

#include "hg_batch_kernels.h"

#ifndef HG_BV
#error "compile with -DHG_BV=<variant>"
#endif
#ifndef HG_RS1
#error "compile with -DHG_RS1=<solver + 1>"
#endif

#define HG_CAT2(a, b) a##b
#define HG_CAT(a, b) HG_CAT2(a, b)

namespace hg {

namespace {
using BvR = std::conditional_t<(HG_BV % 2) == 1, float, double>;
constexpr bool kBvNS = (HG_BV / 2) >= 1;
constexpr bool kBvIP = (HG_BV / 2) == 2;
}  // namespace

#define HG_STAGE_NAME HG_CAT(HG_CAT(HG_CAT(batch_stage_v, HG_BV), _rs), HG_RS1)

hipError_t HG_STAGE_NAME(int tile, const BatchArgs &b, int block, hipStream_t st) {
    return batch_stage_variant<kBvNS, kBvIP, HG_RS1 - 1, BvR, BvR>(tile, b, block, st);
}

hipError_t HG_CAT(HG_STAGE_NAME, _p)(int P, int tile, const BatchArgs &b, hipStream_t st) {
    return batch_stage_p_variant<kBvNS, kBvIP, HG_RS1 - 1, BvR, BvR>(P, tile, b, st);
}

#if HG_RS1 == 0
hipError_t HG_CAT(batch_gradient_v, HG_BV)(int tile, const BatchArgs &b, int block, hipStream_t st) {
    if constexpr (!kBvNS) {
        (void)tile; (void)b; (void)block; (void)st;
        return hipErrorInvalidValue;
    } else {
        return batch_gradient_variant<!kBvIP, BvR, BvR>(tile, b, block, st);
    }
}

hipError_t HG_CAT(HG_CAT(batch_gradient_v, HG_BV), _p)(int P, int tile, const BatchArgs &b, hipStream_t st) {
    if constexpr (!kBvNS) {
        (void)P; (void)tile; (void)b; (void)st;
        return hipErrorInvalidValue;
    } else {
        return batch_gradient_p_variant<!kBvIP, BvR, BvR>(P, tile, b, st);
    }
}
#endif

}  // namespace hg
