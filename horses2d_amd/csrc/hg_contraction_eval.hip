// Measures the volume contraction of the stage kernel (the weak-form derivative
// QDot(i,j,c) = sum_l hatD(i,l) F(l,j,c) + hatD(j,l) G(i,l,c), DGWeakIntegrals.f90:111-181)
// two ways on one MI355X, for P = 7 and 8 in FP64 and FP32:
//   * VALU: one lane per node, as the stage kernel does it (the l loop over LDS operands);
//   * MFMA: the element tile's contraction as two GEMMs per direction on the matrix cores,
//     hatD (P x P, zero-padded to 16 x 4k) times the tile's fluxes arranged as P x (elements 4P)
//     panels (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), results summed through LDS into
//     the node layout the rest of the stage kernel uses.
// Each workgroup holds a tile of EPB elements (the stage kernel's wide tile) in LDS and repeats
// the contraction REPS times, feeding each result back into the flux operands (so neither the
// loads nor the products can be hoisted); the grid covers the GPU many times.  The two results
// are compared (max relative difference) so a wrong operand layout shows as an O(1) error.
//
//   horses2d_amd/bin/contraction_eval            (built by horses2d_amd/build.py)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                                            \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(1);                                                                   \
        }                                                                                   \
    } while (0)

namespace {

constexpr int REPS = 32;

template <int P>
struct Shape {
    static constexpr int PP = P * P;
    static constexpr int EPB = 256 / (4 * P) < 256 / PP ? 256 / (4 * P) : 256 / PP;   // the wide tile
    static constexpr int NN = EPB * PP;
    static constexpr int TPB = ((NN + 63) / 64) * 64;
    static constexpr int NCOL = EPB * 4 * P;             // GEMM columns of the tile: (element, j, c)
    static constexpr int NBLK = (NCOL + 15) / 16;
    static constexpr int KST = (P + 3) / 4;              // k steps of 4
};

// Loads the tile's F and G (node-major, four variables) and hatD into LDS.
template <int P, typename R>
__device__ void load_tile(const R *F, const R *G, const R *H, R *sF, R *sG, R *sH) {
    using S = Shape<P>;
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * S::NN * 4;
    for (int x = t; x < S::NN * 4; x += S::TPB) { sF[x] = F[base + x]; sG[x] = G[base + x]; }
    for (int x = t; x < S::PP; x += S::TPB) sH[x] = H[x];
    __syncthreads();
}

// Feeds the result back into the operands (scaled so the values stay bounded).
template <int P, typename R>
__device__ void feed_back(const R *acc, int t, R *sF, R *sG) {
    using S = Shape<P>;
    if (t < S::NN) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sF[t * 4 + c] = R(0.125) * acc[c] + R(0.5) * sF[t * 4 + c];
            sG[t * 4 + c] = R(0.125) * acc[c] - R(0.5) * sG[t * 4 + c];
        }
    }
}

template <int P, typename R>
__global__ __launch_bounds__(Shape<P>::TPB) void k_valu(const R *F, const R *G, const R *H, R *out) {
    using S = Shape<P>;
    __shared__ R sF[S::NN * 4], sG[S::NN * 4], sH[S::PP];
    load_tile<P, R>(F, G, H, sF, sG, sH);
    const int t = threadIdx.x, s = t / S::PP, n = t % S::PP, i = n % P, j = n / P;
    R acc[4] = {R(0), R(0), R(0), R(0)};
    for (int r = 0; r < REPS; ++r) {
        if (t < S::NN) {
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = R(0);
            const R *cF = sF + s * S::PP * 4, *cG = sG + s * S::PP * 4;
#pragma unroll 1
            for (int l = 0; l < P; ++l) {
                const R hi = sH[i * P + l], hj = sH[j * P + l];
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] = acc[c] + hi * cF[(l + P * j) * 4 + c] + hj * cG[(i + P * l) * 4 + c];
            }
        }
        __syncthreads();
        feed_back<P, R>(acc, t, sF, sG);
        __syncthreads();
    }
    if (t < S::NN)
        for (int c = 0; c < 4; ++c) out[((int64_t)blockIdx.x * S::NN + t) * 4 + c] = acc[c];
}

// The VALU form with each element's node rows padded to P + 1 nodes in LDS: at P = 8 the rows of
// the unpadded layout start 256 B (FP64) or 128 B (FP32) apart, so the lanes of one wavefront that
// read F at the same l and different j hit the same LDS banks.
template <int P, typename R>
__global__ __launch_bounds__(Shape<P>::TPB) void k_valu_pad(const R *F, const R *G, const R *H, R *out) {
    using S = Shape<P>;
    constexpr int PR = P + 1, PE = P * PR;
    __shared__ R sF[S::EPB * PE * 4], sG[S::EPB * PE * 4], sH[S::PP];
    const int t = threadIdx.x, s = t / S::PP, n = t % S::PP, i = n % P, j = n / P;
    const int64_t base = (int64_t)blockIdx.x * S::NN * 4;
    for (int x = t; x < S::NN * 4; x += S::TPB) {
        const int node = x / 4, c = x % 4, e = node / S::PP, m = node % S::PP;
        const int p = (e * PE + m % P + PR * (m / P)) * 4 + c;
        sF[p] = F[base + x];
        sG[p] = G[base + x];
    }
    for (int x = t; x < S::PP; x += S::TPB) sH[x] = H[x];
    __syncthreads();
    const int me = (s * PE + i + PR * j) * 4;
    R acc[4] = {R(0), R(0), R(0), R(0)};
    for (int r = 0; r < REPS; ++r) {
        if (t < S::NN) {
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = R(0);
            const R *cF = sF + s * PE * 4, *cG = sG + s * PE * 4;
#pragma unroll 1
            for (int l = 0; l < P; ++l) {
                const R hi = sH[i * P + l], hj = sH[j * P + l];
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] = acc[c] + hi * cF[(l + PR * j) * 4 + c] + hj * cG[(i + PR * l) * 4 + c];
            }
        }
        __syncthreads();
        if (t < S::NN) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                sF[me + c] = R(0.125) * acc[c] + R(0.5) * sF[me + c];
                sG[me + c] = R(0.125) * acc[c] - R(0.5) * sG[me + c];
            }
        }
        __syncthreads();
    }
    if (t < S::NN)
        for (int c = 0; c < 4; ++c) out[((int64_t)blockIdx.x * S::NN + t) * 4 + c] = acc[c];
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
__device__ inline v4d mfma(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline v4f mfma(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
template <typename R>
struct Vec4;
template <>
struct Vec4<double> { using T = v4d; };
template <>
struct Vec4<float> { using T = v4f; };

// One direction's GEMM for the column blocks of this wavefront: D[row][col] = sum_k hatD[row][k]
// B[k][col], with B[k][(e, a, c)] the flux of element e at the node whose index along the
// contracted direction is k and the other index is a.  Operand layout of the 16x16x4 MFMA:
// lane L supplies A[L % 16][L / 16] and B[L / 16][L % 16]; it receives D[4 (L / 16) + r][L % 16]
// (FP32: four consecutive rows per lane group) or D[4 r + L / 16][L % 16] (FP64: the rows
// interleaved over the lane groups).
template <int P, typename R, bool XI>
__device__ void gemm_dir(const R *sH, const R *sB, R *sOut, bool add) {
    using S = Shape<P>;
    using V = typename Vec4<R>::T;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = S::TPB / 64;
    const int row = lane & 15, kq = lane >> 4;
    for (int blk = wave; blk < S::NBLK; blk += nw) {
        const int col = blk * 16 + (lane & 15);
        const bool cvalid = col < S::NCOL;
        const int e = col / (4 * P), rem = col % (4 * P), a = rem / 4, c = rem % 4;
        V d = {R(0), R(0), R(0), R(0)};
#pragma unroll
        for (int ks = 0; ks < S::KST; ++ks) {
            const int k = ks * 4 + kq;
            const R av = (row < P && k < P) ? sH[row * P + k] : R(0);
            R bv = R(0);
            if (cvalid && k < P) {
                // XI: F(l = k, j = a);  eta: G(i = a, l = k)
                const int node = XI ? (k + P * a) : (a + P * k);
                bv = sB[(e * S::PP + node) * 4 + c];
            }
            d = mfma(av, bv, d);
        }
        // D rows 4 kq + r (the free index of the output), column (e, a, c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int orow = sizeof(R) == 8 ? 4 * r + kq : 4 * kq + r;
            if (cvalid && orow < P) {
                // XI: row i, a = j;  eta: row j, a = i
                const int node = XI ? (orow + P * a) : (a + P * orow);
                const int o = (e * S::PP + node) * 4 + c;
                const R v = d[r];
                sOut[o] = add ? sOut[o] + v : v;
            }
        }
    }
}

template <int P, typename R>
__global__ __launch_bounds__(Shape<P>::TPB) void k_mfma(const R *F, const R *G, const R *H, R *out) {
    using S = Shape<P>;
    __shared__ R sF[S::NN * 4], sG[S::NN * 4], sH[S::PP], sO[S::NN * 4];
    load_tile<P, R>(F, G, H, sF, sG, sH);
    const int t = threadIdx.x;
    R acc[4] = {R(0), R(0), R(0), R(0)};
    for (int r = 0; r < REPS; ++r) {
        gemm_dir<P, R, true>(sH, sF, sO, false);
        __syncthreads();
        gemm_dir<P, R, false>(sH, sG, sO, true);
        __syncthreads();
        if (t < S::NN) {
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = sO[t * 4 + c];
        }
        feed_back<P, R>(acc, t, sF, sG);
        __syncthreads();
    }
    if (t < S::NN)
        for (int c = 0; c < 4; ++c) out[((int64_t)blockIdx.x * S::NN + t) * 4 + c] = acc[c];
}

template <int P, typename R>
void run(const char *prec) {
    using S = Shape<P>;
    const int blocks = 256 * 8 * 6;   // many rounds of resident workgroups
    const size_t n = (size_t)blocks * S::NN * 4;
    std::vector<R> hF(n), hG(n), hH(S::PP);
    uint64_t x = 88172645463325252ull;
    auto rnd = [&]() { x ^= x << 13; x ^= x >> 7; x ^= x << 17; return R((double)(x >> 11) / 9007199254740992.0 - 0.5); };
    for (auto &v : hF) v = rnd();
    for (auto &v : hG) v = rnd();
    for (auto &v : hH) v = rnd();
    R *F, *G, *H, *o1, *o2, *o3;
    CHECK(hipMalloc(&F, n * sizeof(R)));
    CHECK(hipMalloc(&G, n * sizeof(R)));
    CHECK(hipMalloc(&H, S::PP * sizeof(R)));
    CHECK(hipMalloc(&o1, n * sizeof(R)));
    CHECK(hipMalloc(&o2, n * sizeof(R)));
    CHECK(hipMalloc(&o3, n * sizeof(R)));
    CHECK(hipMemcpy(F, hF.data(), n * sizeof(R), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(G, hG.data(), n * sizeof(R), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(H, hH.data(), S::PP * sizeof(R), hipMemcpyHostToDevice));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    float ms[3] = {0, 0, 0};
    for (int v = 0; v < 3; ++v) {
        for (int it = 0; it < 4; ++it) {   // warm-up, then the timed launch
            if (it == 3) CHECK(hipEventRecord(a));
            if (v == 0) hipLaunchKernelGGL((k_valu<P, R>), dim3(blocks), dim3(S::TPB), 0, 0, F, G, H, o1);
            else if (v == 1) hipLaunchKernelGGL((k_mfma<P, R>), dim3(blocks), dim3(S::TPB), 0, 0, F, G, H, o2);
            else hipLaunchKernelGGL((k_valu_pad<P, R>), dim3(blocks), dim3(S::TPB), 0, 0, F, G, H, o3);
            CHECK(hipGetLastError());
            if (it == 3) CHECK(hipEventRecord(b));
        }
        CHECK(hipEventSynchronize(b));
        CHECK(hipEventElapsedTime(&ms[v], a, b));
    }
    std::vector<R> r1(n), r2(n), r3(n);
    CHECK(hipMemcpy(r1.data(), o1, n * sizeof(R), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(r2.data(), o2, n * sizeof(R), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(r3.data(), o3, n * sizeof(R), hipMemcpyDeviceToHost));
    double scale = 0.0, diff = 0.0, diff3 = 0.0;
    for (size_t k = 0; k < n; ++k) {
        scale = std::fmax(scale, std::fabs((double)r1[k]));
        diff = std::fmax(diff, std::fabs((double)r1[k] - (double)r2[k]));
        diff3 = std::fmax(diff3, std::fabs((double)r1[k] - (double)r3[k]));
    }
    // useful multiply-adds: two directions, 4 variables, P^3 per element
    const double macs = (double)blocks * S::EPB * 2.0 * 4.0 * P * P * P * REPS;
    const char *name[3] = {"VALU", "MFMA", "VALU, rows padded"};
    for (int v = 0; v < 3; ++v)
        std::printf("P=%d %s %-18s EPB=%d TPB=%d  %8.3f ms  %7.2f us/tile-rep-per-CU  useful %7.2f TFLOP/s\n", P, prec,
                    name[v], S::EPB, S::TPB, ms[v], 1e3 * ms[v] / ((double)blocks * REPS / 256.0),
                    2.0 * macs / (ms[v] * 1e-3) / 1e12);
    std::printf("P=%d %s time / VALU: MFMA %.3f, padded %.3f; max rel difference to VALU: MFMA %.2e, padded %.2e\n",
                P, prec, ms[1] / ms[0], ms[2] / ms[0], diff / scale, diff3 / scale);
    CHECK(hipFree(F)); CHECK(hipFree(G)); CHECK(hipFree(H)); CHECK(hipFree(o1)); CHECK(hipFree(o2)); CHECK(hipFree(o3));
}

}  // namespace

int main() {
    run<6, double>("fp64");
    run<7, double>("fp64");
    run<8, double>("fp64");
    run<9, double>("fp64");
    run<6, float>("fp32");
    run<7, float>("fp32");
    run<8, float>("fp32");
    run<9, float>("fp32");
    return 0;
}
