// An entry point that allocates nothing, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// hg_conservation_sums_noalloc computes what hg_conservation_sums does, with the same kernels
// and the same summation order, but keeps its per-workgroup partial sums in a file-scope device
// array of fixed size instead of the handle's buffer, writes the four sums to a device pointer
// the caller passes and does not synchronise, so it can be captured into a graph.  Every call on
// a GPU shares that GPU's array (each GPU holds its own copy until the process ends): calls that
// overlap on one GPU race.  Every partial a call reads is written by the same call first.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "hg_device.h"
#include "hg_handle.h"
#include "horses2d_gpu.h"

namespace {

// 2 MiB of partial sums: four per workgroup of the conservation-sum kernel, enough for the
// 512 x 512 order-7 cylinder grid (65536 workgroups)
constexpr size_t kScratchDoubles = 262144;
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mutex;
std::vector<double *> g_scratch;   // per device: the address of that device's copy of `scratch`

}  // namespace

namespace hg {

// Finds this device's copy of the scratch array (once per device, with that device current).
// Called by hg_create, so that the entry point itself never queries the runtime.
int noalloc_setup(int device) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (device < 0) return handle_error("hg: bad device");
    if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
    if (g_scratch[device]) return 0;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    double *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != device) hipSetDevice(cur);
    if (e != hipSuccess || !p)
        return handle_error(std::string("hg: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return 0;
}

}  // namespace hg

extern "C" int hg_conservation_sums_noalloc(hg_handle *h, double *out_dev) {
    if (!h || !out_dev) return hg::handle_error("hg_conservation_sums_noalloc: null argument");
    if (!h->sub.empty())
        return hg::handle_error("hg_conservation_sums_noalloc: one polynomial order per mesh (use hg_conservation_sums)");
    if ((size_t)h->cons_nb * 4 > kScratchDoubles)
        return hg::handle_error("hg_conservation_sums_noalloc: the mesh needs " + std::to_string(h->cons_nb) +
                                " partial sums, above the scratch array's " + std::to_string(kScratchDoubles / 4));
    double *scr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (h->device >= 0 && h->device < (int)g_scratch.size()) scr = g_scratch[h->device];
    }
    if (!scr) return hg::handle_error("hg_conservation_sums_noalloc: scratch array not set up for this device");
    hipError_t e = hg::launch_cons_sums(h->P, h->Q, h->jac64, h->w, h->K, scr, h->cons_nb, out_dev, h->stream);
    if (e != hipSuccess)
        return hg::handle_error(std::string("hg_conservation_sums_noalloc: ") + hipGetErrorString(e));
    return 0;
}
