// Batched form: one launch per RK stage (and per gradient pass) over a work list of element
// tiles drawn from many independent cases of different mesh size and polynomial order.
// Workgroup k evaluates work item sched[k] with the same stage_body / gradient_body as the
// single-case kernels, so the per-element arithmetic, and therefore every result bit, is the
// one of the single-case path.  The LDS of the workgroup is the union of the per-order tile
// layouts.  A grid-stride (persistent) loop over the list was measured to raise the register
// allocation from 149 to 213-256 VGPRs (occupancy 3 -> 2 or 1 wavefronts per SIMD): the
// compiler keeps every order's loop-invariant lane index math live across iterations.  So
// the launch has one workgroup per work item and the hardware dispatcher walks the list in
// the order sched gives it.
#pragma once
#include "hg_kernels.h"

namespace hg {

template <int P, int TILE>
struct TileOf {
    static constexpr int EPB = TILE == 0 ? Tiles<P>::small : Tiles<P>::large;
};

template <int TILE, bool NS, bool IP, typename R>
union BatchStageSmem {
    StageSmem<2, TileOf<2, TILE>::EPB, NS, IP, R> p2;
    StageSmem<3, TileOf<3, TILE>::EPB, NS, IP, R> p3;
    StageSmem<4, TileOf<4, TILE>::EPB, NS, IP, R> p4;
    StageSmem<5, TileOf<5, TILE>::EPB, NS, IP, R> p5;
    StageSmem<6, TileOf<6, TILE>::EPB, NS, IP, R> p6;
    StageSmem<7, TileOf<7, TILE>::EPB, NS, IP, R> p7;
    StageSmem<8, TileOf<8, TILE>::EPB, NS, IP, R> p8;
    StageSmem<9, TileOf<9, TILE>::EPB, NS, IP, R> p9;
    StageSmem<10, TileOf<10, TILE>::EPB, NS, IP, R> p10;
    StageSmem<11, TileOf<11, TILE>::EPB, NS, IP, R> p11;
};

template <int TILE, bool BR1, typename R>
union BatchGradSmem {
    GradSmem<2, TileOf<2, TILE>::EPB, BR1, R> p2;
    GradSmem<3, TileOf<3, TILE>::EPB, BR1, R> p3;
    GradSmem<4, TileOf<4, TILE>::EPB, BR1, R> p4;
    GradSmem<5, TileOf<5, TILE>::EPB, BR1, R> p5;
    GradSmem<6, TileOf<6, TILE>::EPB, BR1, R> p6;
    GradSmem<7, TileOf<7, TILE>::EPB, BR1, R> p7;
    GradSmem<8, TileOf<8, TILE>::EPB, BR1, R> p8;
    GradSmem<9, TileOf<9, TILE>::EPB, BR1, R> p9;
    GradSmem<10, TileOf<10, TILE>::EPB, BR1, R> p10;
    GradSmem<11, TileOf<11, TILE>::EPB, BR1, R> p11;
};

#define HG_BATCH_CASE(PV, CALL) \
    case PV: CALL(PV, p##PV); break;

// The work item is uniform across the workgroup: read it into scalar registers, so that the
// case's StageArgs is addressed uniformly and its fields load through the scalar cache.
__device__ inline int4 uniform_item(const int4 *__restrict__ items, const int *__restrict__ sched, int k) {
    const int idx = __builtin_amdgcn_readfirstlane(sched[k]);
    const int4 v = items[idx];
    return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                     __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}

template <int TILE, bool NS, bool IP, int RS, typename R, typename TR>
__global__ __launch_bounds__(256) void k_batch_stage(const StageArgs *__restrict__ slots, const int4 *__restrict__ items,
                                                     const int *__restrict__ sched, const double *__restrict__ dt,
                                                     BatchArgs b) {
    __shared__ BatchStageSmem<TILE, NS, IP, R> S;
    const int k = xcd_tile(blockIdx.x, gridDim.x);   // position in the schedule
    if (k < b.nitems) {
        const int4 it = uniform_item(items, sched, k);
        const StageArgs &a = slots[(int64_t)it.x * b.nslots + b.slot];
        double cq = a.cq, cg = a.cg;
        if (dt) {
            const double d = dt[it.x];
            if (b.dt_cq) cq *= d;
            if (b.dt_cg) cg *= d;
        }
#define HG_CALL(PV, M) stage_body<PV, TileOf<PV, TILE>::EPB, NS, IP, RS, R, TR>(a, it.y, it.w, cq, cg, S.M)
        switch (it.z) {
            HG_BATCH_CASE(2, HG_CALL)
            HG_BATCH_CASE(3, HG_CALL)
            HG_BATCH_CASE(4, HG_CALL)
            HG_BATCH_CASE(5, HG_CALL)
            HG_BATCH_CASE(6, HG_CALL)
            HG_BATCH_CASE(7, HG_CALL)
            HG_BATCH_CASE(8, HG_CALL)
            HG_BATCH_CASE(9, HG_CALL)
            HG_BATCH_CASE(10, HG_CALL)
            HG_BATCH_CASE(11, HG_CALL)
            default: break;
        }
#undef HG_CALL
    }
}

template <int TILE, bool BR1, typename R, typename TR>
__global__ __launch_bounds__(256) void k_batch_gradient(const StageArgs *__restrict__ slots,
                                                        const int4 *__restrict__ items,
                                                        const int *__restrict__ sched, BatchArgs b) {
    __shared__ BatchGradSmem<TILE, BR1, R> S;
    const int k = xcd_tile(blockIdx.x, gridDim.x);   // position in the schedule
    if (k < b.nitems) {
        const int4 it = uniform_item(items, sched, k);
        const StageArgs &a = slots[(int64_t)it.x * b.nslots + b.slot];
#define HG_CALL(PV, M) gradient_body<PV, TileOf<PV, TILE>::EPB, BR1, R, TR, true>(a, it.y, S.M)
        switch (it.z) {
            HG_BATCH_CASE(2, HG_CALL)
            HG_BATCH_CASE(3, HG_CALL)
            HG_BATCH_CASE(4, HG_CALL)
            HG_BATCH_CASE(5, HG_CALL)
            HG_BATCH_CASE(6, HG_CALL)
            HG_BATCH_CASE(7, HG_CALL)
            HG_BATCH_CASE(8, HG_CALL)
            HG_BATCH_CASE(9, HG_CALL)
            HG_BATCH_CASE(10, HG_CALL)
            HG_BATCH_CASE(11, HG_CALL)
            default: break;
        }
#undef HG_CALL
    }
}

#undef HG_BATCH_CASE

// ---- persistent form: one launch per RK stage whose workgroups stay resident and loop.  The
// grid is what the GPU holds at once; the host splits it over the orders in proportion to their
// work, in multiples of eight (pg0 / pgn, filled per launch).  Workgroup r of order P (on XCD
// r & 7) takes every (pgn / 8)-th item of XCD r & 7's contiguous eighth of P's list, so that the
// traces a tile reads from its neighbours were mostly written through the same L2.  The loop
// sits inside each order's case of the switch: a workgroup keeps one order, and the registers
// are those of the costliest order, as in the one-item-per-workgroup kernel.
struct PersistGrid {
    int pg0[12], pgn[12];
};

template <int P, int TILE, bool NS, bool IP, int RS, typename R, typename TR, typename Smem>
__device__ __forceinline__ void persist_stage(const StageArgs *__restrict__ slots, const int4 *__restrict__ items,
                                              const double *__restrict__ dt, const BatchArgs &b, int pgn, int r,
                                              Smem &S) {
    constexpr int EPB = TileOf<P, TILE>::EPB;
    const int n = b.pn[P], per = pgn >> 3, x = r & 7;
    const int cs = (int)((int64_t)n * x / 8), ce = (int)((int64_t)n * (x + 1) / 8);
    for (int k = cs + (r >> 3); k < ce; k += per) {
        if (k != cs + (r >> 3)) __syncthreads();   // the previous item's LDS is read
        const int4 it = uniform_item(items, b.psched[P], k);
        const StageArgs &a = slots[(int64_t)it.x * b.nslots + b.slot];
        double cq = a.cq, cg = a.cg;
        if (dt) {
            const double d = dt[it.x];
            if (b.dt_cq) cq *= d;
            if (b.dt_cg) cg *= d;
        }
        stage_body<P, EPB, NS, IP, RS, R, TR>(a, it.y, it.w, cq, cg, S);
    }
}

template <int P, int TILE, bool BR1, typename R, typename TR, typename Smem>
__device__ __forceinline__ void persist_gradient(const StageArgs *__restrict__ slots, const int4 *__restrict__ items,
                                                 const BatchArgs &b, int pgn, int r, Smem &S) {
    constexpr int EPB = TileOf<P, TILE>::EPB;
    const int n = b.pn[P], per = pgn >> 3, x = r & 7;
    const int cs = (int)((int64_t)n * x / 8), ce = (int)((int64_t)n * (x + 1) / 8);
    for (int k = cs + (r >> 3); k < ce; k += per) {
        if (k != cs + (r >> 3)) __syncthreads();
        const int4 it = uniform_item(items, b.psched[P], k);
        const StageArgs &a = slots[(int64_t)it.x * b.nslots + b.slot];
        gradient_body<P, EPB, BR1, R, TR, true>(a, it.y, S);
    }
}

// The order of workgroup w and its rank among that order's workgroups (uniform).
__device__ inline int persist_order(const PersistGrid &g, int w, int &r) {
    int P = 0;
    for (int p = 2; p < 12; ++p)
        if (w >= g.pg0[p] && w < g.pg0[p] + g.pgn[p]) P = p;
    r = w - g.pg0[P];
    return P;
}

#define HG_PERSIST_CASE(PV, CALL) \
    case PV: CALL(PV, p##PV); break;

// The FP32 form keeps the occupancy of the one-item kernel (eight wavefronts per SIMD, 64
// VGPRs): without the bound the loop's hoisted lane-index values take it to 86 VGPRs.
template <typename R>
struct PersistWaves {
    static constexpr int value = sizeof(R) == 4 ? 8 : 1;
};

template <int TILE, bool NS, bool IP, int RS, typename R, typename TR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PersistWaves<R>::value))) void k_batch_stage_persist(const StageArgs *__restrict__ slots,
                                                             const int4 *__restrict__ items,
                                                             const double *__restrict__ dt, BatchArgs b,
                                                             PersistGrid g) {
    __shared__ BatchStageSmem<TILE, NS, IP, R> S;
    int r = 0;
    const int P = persist_order(g, blockIdx.x, r);
#define HG_CALL(PV, M) persist_stage<PV, TILE, NS, IP, RS, R, TR>(slots, items, dt, b, g.pgn[PV], r, S.M)
    switch (P) {
        HG_PERSIST_CASE(2, HG_CALL)
        HG_PERSIST_CASE(3, HG_CALL)
        HG_PERSIST_CASE(4, HG_CALL)
        HG_PERSIST_CASE(5, HG_CALL)
        HG_PERSIST_CASE(6, HG_CALL)
        HG_PERSIST_CASE(7, HG_CALL)
        HG_PERSIST_CASE(8, HG_CALL)
        HG_PERSIST_CASE(9, HG_CALL)
        HG_PERSIST_CASE(10, HG_CALL)
        HG_PERSIST_CASE(11, HG_CALL)
        default: break;
    }
#undef HG_CALL
}

template <int TILE, bool BR1, typename R, typename TR>
__global__ __launch_bounds__(256) void k_batch_gradient_persist(const StageArgs *__restrict__ slots,
                                                                const int4 *__restrict__ items, BatchArgs b,
                                                                PersistGrid g) {
    __shared__ BatchGradSmem<TILE, BR1, R> S;
    int r = 0;
    const int P = persist_order(g, blockIdx.x, r);
#define HG_CALL(PV, M) persist_gradient<PV, TILE, BR1, R, TR>(slots, items, b, g.pgn[PV], r, S.M)
    switch (P) {
        HG_PERSIST_CASE(2, HG_CALL)
        HG_PERSIST_CASE(3, HG_CALL)
        HG_PERSIST_CASE(4, HG_CALL)
        HG_PERSIST_CASE(5, HG_CALL)
        HG_PERSIST_CASE(6, HG_CALL)
        HG_PERSIST_CASE(7, HG_CALL)
        HG_PERSIST_CASE(8, HG_CALL)
        HG_PERSIST_CASE(9, HG_CALL)
        HG_PERSIST_CASE(10, HG_CALL)
        HG_PERSIST_CASE(11, HG_CALL)
        default: break;
    }
#undef HG_CALL
}
#undef HG_PERSIST_CASE

// The persistent grid of a kernel: the resident workgroups of the whole GPU (occupancy times
// compute units, once per kernel), split over the orders of b.pn in proportion to their nodes
// (items x elements per item x P^2: the kernels are bound by memory traffic, which follows the
// nodes; a P^3 weighting was slower), in multiples of eight, at least eight per order with work.
template <int TILE, typename K>
hipError_t persist_grid(K kernel, int block, const BatchArgs &b, PersistGrid &g, int &grid) {
    static int cap = 0;
    if (!cap) {
        int dev = 0, cus = 0, per = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0);
        if (e != hipSuccess) return e;
        cap = std::max(8, cus * std::max(per, 1));
    }
    double work[12] = {0}, tot = 0.0;
    for (int p = 2; p < 12; ++p) {
        const int epb = TILE == 0 ? tile_elems(p, 0) : tile_elems(p, 1);
        work[p] = (double)b.pn[p] * epb * p * p;
        tot += work[p];
    }
    grid = 0;
    for (int p = 0; p < 12; ++p) {
        g.pg0[p] = grid;
        g.pgn[p] = 0;
        if (p < 2 || b.pn[p] <= 0) continue;
        int n8 = (int)(cap / 8 * work[p] / tot + 0.5);
        n8 = std::max(1, std::min(n8, (b.pn[p] + 7) / 8));   // no more workgroups than items
        g.pgn[p] = 8 * n8;
        grid += g.pgn[p];
    }
    return hipSuccess;
}

// One order per launch: the work items of the cases of order P only, with that order's own LDS
// tile and register allocation (the union kernels above reserve the largest order's LDS and the
// register count of the costliest order for every item).  Same stage_body / gradient_body, so
// the same bits.
template <int P, int TILE, bool NS, bool IP, int RS, typename R, typename TR>
__global__ __launch_bounds__((Tile<P, TileOf<P, TILE>::EPB>::TPB)) void k_batch_stage_p(
    const StageArgs *__restrict__ slots, const int4 *__restrict__ items, const int *__restrict__ sched,
    const double *__restrict__ dt, BatchArgs b) {
    constexpr int EPB = TileOf<P, TILE>::EPB;
    __shared__ StageSmem<P, EPB, NS, IP, R> S;
    const int k = xcd_tile(blockIdx.x, gridDim.x);
    if (k < b.nitems) {
        const int4 it = uniform_item(items, sched, k);
        const StageArgs &a = slots[(int64_t)it.x * b.nslots + b.slot];
        double cq = a.cq, cg = a.cg;
        if (dt) {
            const double d = dt[it.x];
            if (b.dt_cq) cq *= d;
            if (b.dt_cg) cg *= d;
        }
        stage_body<P, EPB, NS, IP, RS, R, TR>(a, it.y, it.w, cq, cg, S);
    }
}

template <int P, int TILE, bool BR1, typename R, typename TR>
__global__ __launch_bounds__((Tile<P, TileOf<P, TILE>::EPB>::TPB)) void k_batch_gradient_p(
    const StageArgs *__restrict__ slots, const int4 *__restrict__ items, const int *__restrict__ sched, BatchArgs b) {
    constexpr int EPB = TileOf<P, TILE>::EPB;
    __shared__ GradSmem<P, EPB, BR1, R> S;
    const int k = xcd_tile(blockIdx.x, gridDim.x);
    if (k < b.nitems) {
        const int4 it = uniform_item(items, sched, k);
        const StageArgs &a = slots[(int64_t)it.x * b.nslots + b.slot];
        gradient_body<P, EPB, BR1, R, TR, true>(a, it.y, S);
    }
}

template <int P, bool NS, bool IP, int RS, typename R, typename TR>
hipError_t batch_stage_p_launch(int tile, const BatchArgs &b, hipStream_t st) {
    if (tile == 0)
        hipLaunchKernelGGL((k_batch_stage_p<P, 0, NS, IP, RS, R, TR>), dim3(b.nitems),
                           dim3(Tile<P, TileOf<P, 0>::EPB>::TPB), 0, st, b.slots, b.items, b.sched, b.dt, b);
    else
        hipLaunchKernelGGL((k_batch_stage_p<P, 1, NS, IP, RS, R, TR>), dim3(b.nitems),
                           dim3(Tile<P, TileOf<P, 1>::EPB>::TPB), 0, st, b.slots, b.items, b.sched, b.dt, b);
    return hipGetLastError();
}

template <int P, bool BR1, typename R, typename TR>
hipError_t batch_gradient_p_launch(int tile, const BatchArgs &b, hipStream_t st) {
    if (tile == 0)
        hipLaunchKernelGGL((k_batch_gradient_p<P, 0, BR1, R, TR>), dim3(b.nitems),
                           dim3(Tile<P, TileOf<P, 0>::EPB>::TPB), 0, st, b.slots, b.items, b.sched, b);
    else
        hipLaunchKernelGGL((k_batch_gradient_p<P, 1, BR1, R, TR>), dim3(b.nitems),
                           dim3(Tile<P, TileOf<P, 1>::EPB>::TPB), 0, st, b.slots, b.items, b.sched, b);
    return hipGetLastError();
}

#define HG_BATCH_ORDERS(CALL) \
    switch (P) {                  \
    case 2: return CALL(2);       \
    case 3: return CALL(3);       \
    case 4: return CALL(4);       \
    case 5: return CALL(5);       \
    case 6: return CALL(6);       \
    case 7: return CALL(7);       \
    case 8: return CALL(8);       \
    case 9: return CALL(9);       \
    case 10: return CALL(10);     \
    case 11: return CALL(11);     \
    default: return hipErrorInvalidValue; \
    }

template <bool NS, bool IP, int RS, typename R, typename TR>
hipError_t batch_stage_p_variant(int P, int tile, const BatchArgs &b, hipStream_t st) {
#define HG_CALL(PV) batch_stage_p_launch<PV, NS, IP, RS, R, TR>(tile, b, st)
    HG_BATCH_ORDERS(HG_CALL)
#undef HG_CALL
}

template <bool BR1, typename R, typename TR>
hipError_t batch_gradient_p_variant(int P, int tile, const BatchArgs &b, hipStream_t st) {
#define HG_CALL(PV) batch_gradient_p_launch<PV, BR1, R, TR>(tile, b, st)
    HG_BATCH_ORDERS(HG_CALL)
#undef HG_CALL
}
#undef HG_BATCH_ORDERS

// Launchers of one physics/precision variant (instantiated by hg_batch_inst.hip); grid = nitems.
template <bool NS, bool IP, int RS, typename R, typename TR>
hipError_t batch_stage_variant(int tile, const BatchArgs &b, int block, hipStream_t st) {
    if (b.persist) {
        PersistGrid g{};
        int grid = 0;
        hipError_t e = tile == 0 ? persist_grid<0>(k_batch_stage_persist<0, NS, IP, RS, R, TR>, block, b, g, grid)
                                 : persist_grid<1>(k_batch_stage_persist<1, NS, IP, RS, R, TR>, block, b, g, grid);
        if (e != hipSuccess) return e;
        if (grid == 0) return hipSuccess;
        if (tile == 0)
            hipLaunchKernelGGL((k_batch_stage_persist<0, NS, IP, RS, R, TR>), dim3(grid), dim3(block), 0, st, b.slots,
                               b.items, b.dt, b, g);
        else
            hipLaunchKernelGGL((k_batch_stage_persist<1, NS, IP, RS, R, TR>), dim3(grid), dim3(block), 0, st, b.slots,
                               b.items, b.dt, b, g);
        return hipGetLastError();
    }
    if (tile == 0)
        hipLaunchKernelGGL((k_batch_stage<0, NS, IP, RS, R, TR>), dim3(b.nitems), dim3(block), 0, st, b.slots, b.items,
                           b.sched, b.dt, b);
    else
        hipLaunchKernelGGL((k_batch_stage<1, NS, IP, RS, R, TR>), dim3(b.nitems), dim3(block), 0, st, b.slots, b.items,
                           b.sched, b.dt, b);
    return hipGetLastError();
}

template <bool BR1, typename R, typename TR>
hipError_t batch_gradient_variant(int tile, const BatchArgs &b, int block, hipStream_t st) {
    if (b.persist) {
        PersistGrid g{};
        int grid = 0;
        hipError_t e = tile == 0 ? persist_grid<0>(k_batch_gradient_persist<0, BR1, R, TR>, block, b, g, grid)
                                 : persist_grid<1>(k_batch_gradient_persist<1, BR1, R, TR>, block, b, g, grid);
        if (e != hipSuccess) return e;
        if (grid == 0) return hipSuccess;
        if (tile == 0)
            hipLaunchKernelGGL((k_batch_gradient_persist<0, BR1, R, TR>), dim3(grid), dim3(block), 0, st, b.slots,
                               b.items, b, g);
        else
            hipLaunchKernelGGL((k_batch_gradient_persist<1, BR1, R, TR>), dim3(grid), dim3(block), 0, st, b.slots,
                               b.items, b, g);
        return hipGetLastError();
    }
    if (tile == 0)
        hipLaunchKernelGGL((k_batch_gradient<0, BR1, R, TR>), dim3(b.nitems), dim3(block), 0, st, b.slots, b.items,
                           b.sched, b);
    else
        hipLaunchKernelGGL((k_batch_gradient<1, BR1, R, TR>), dim3(b.nitems), dim3(block), 0, st, b.slots, b.items,
                           b.sched, b);
    return hipGetLastError();
}

}  // namespace hg
