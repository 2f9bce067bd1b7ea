// One polynomial order's kernels: compiled once per order with -DHG_P=<N+1> (2..11) so
// that the orders build in parallel and each object holds only its own instantiations.
#include "hg_kernels.h"

#ifndef HG_P
#error "compile with -DHG_P=<points per direction>"
#endif

namespace hg {
HG_INSTANTIATE_ORDER(HG_P)
}  // namespace hg
