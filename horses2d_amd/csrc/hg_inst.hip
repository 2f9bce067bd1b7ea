// One polynomial order's kernels for one set of equations: compiled once per order and
// physics with -DHG_P=<N+1> (2..13) and -DHG_PHYS=<0 Euler | 1 Navier-Stokes BR1 | 2 IP>, so
// that the objects build in parallel.  The order's gradient, trace and reduction kernels
// live in the HG_PHYS=0 object.
#include "hg_kernels.h"

#ifndef HG_P
#error "compile with -DHG_P=<points per direction>"
#endif
#ifndef HG_PHYS
#error "compile with -DHG_PHYS=<0|1|2>"
#endif

namespace hg {
HG_INSTANTIATE_STAGE(HG_P, HG_PHYS)
#if HG_PHYS == 0
HG_INSTANTIATE_ORDER(HG_P)
#endif
}  // namespace hg
