// Batched form: many independent cases advanced together.
// First version: per-stage launches of every case on the batch's own stream.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "horses2d_gpu.h"

struct hg_batch {
    std::vector<hg_handle *> cases;
    hg_batch_options opt{};
};

extern "C" {

int hg_batch_create(hg_handle *const *cases, int32_t ncases, const hg_batch_options *opt, hg_batch **out) {
    if (!cases || ncases <= 0 || !out) return 1;
    hg_batch *b = new hg_batch();
    b->cases.assign(cases, cases + ncases);
    if (opt) b->opt = *opt;
    *out = b;
    return 0;
}

int hg_batch_rk_step(hg_batch *b, const double *t, const double *dt) {
    if (!b || !t || !dt) return 1;
    for (size_t c = 0; c < b->cases.size(); ++c)
        if (hg_rk_step(b->cases[c], t[c], dt[c])) return 1;
    return 0;
}

int hg_batch_synchronize(hg_batch *b) {
    if (!b) return 1;
    for (auto *h : b->cases)
        if (hg_synchronize(h)) return 1;
    return 0;
}

int hg_batch_destroy(hg_batch *b) {
    delete b;
    return 0;
}

}  // extern "C"
