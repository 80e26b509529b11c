// Point transform of the pipeline and device-side timing hooks for the
// benchmark (hipEvents on the tree's stream).
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.hpp"

namespace prx {
void host_unit_transform(const prx_point* p, size_t n, double scale, double m[16], double mi[16]);
void host_apply_transform(const prx_point* in, size_t n, const double m[16], prx_point* out);
int time_operator(prx_tree* t, int precision, int reps, double* ms_gather, double* ms_stencil, int op);

// Streaming copy: one 16-byte vector per thread over a flat grid,
// non-temporal loads and stores (the fastest copy shape measured on gfx950).
__global__ __launch_bounds__(256) void k_copy16(const float* __restrict__ a, float* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st_nt(b + 4 * i, ld_nt(a + 4 * i));
}

// The tile kernels' memory access pattern without their arithmetic: each
// 256-thread workgroup copies one 4096-node brick from each of S input
// arrays to each of S output arrays (16-byte loads, then a barrier, then
// non-temporal 16-byte stores), workgroups mapped to the XCDs in chunks of
// consecutive bricks as the tile kernels map them.  Its bandwidth is the
// ceiling of a pass that reads and writes S vectors brick by brick.
template <typename T>
struct BrickPtrs {
    const T* in[3];
    T* out[3];
};
template <typename T, int S>
__global__ __launch_bounds__(256) void k_brick_copy(BrickPtrs<T> p, uint32_t chunk) {
    using NV = typename NV16<T>::type;
    constexpr int VEC = 16 / (int)sizeof(T), NV_PER = 4096 / (256 * VEC);
    uint32_t bid = blockIdx.x;
    if (chunk) {
        const uint32_t j = bid >> 3;
        bid = (j / chunk) * 8u * chunk + (bid & 7u) * chunk + j % chunk;
    }
    const uint64_t base = (uint64_t)bid * 4096;
    NV v[S][NV_PER];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int k = 0; k < NV_PER; ++k) v[s][k] = ld_v(p.in[s] + base + (uint64_t)(k * 256 + threadIdx.x) * VEC);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int k = 0; k < NV_PER; ++k) st_nt(p.out[s] + base + (uint64_t)(k * 256 + threadIdx.x) * VEC, v[s][k]);
}
}  // namespace prx

extern "C" {

int prx_unit_transform(const prx_point* points, size_t n, double scale, double* m, double* minv, prx_point* out) {
    if (n > 0 && !points) return PRX_ERR_INVALID;
    double mm[16], mi[16];
    prx::host_unit_transform(points, n, scale, mm, mi);
    if (m) std::memcpy(m, mm, sizeof mm);
    if (minv) std::memcpy(minv, mi, sizeof mi);
    if (out) prx::host_apply_transform(points, n, mm, out);
    return PRX_OK;
}

int prx_time_operator(prx_tree* t, int precision, int reps, double* ms_gather, double* ms_stencil) {
    if (!t || reps <= 0) return PRX_ERR_INVALID;
    return prx::time_operator(t, precision, reps, ms_gather, ms_stencil, prx::OP_APPLY);
}

int prx_time_stencil_op(prx_tree* t, int precision, int op, int reps, double* ms_gather, double* ms_stencil) {
    if (!t || reps <= 0 || op < 0 || op > 3) return PRX_ERR_INVALID;
    return prx::time_operator(t, precision, reps, ms_gather, ms_stencil,
                              op == 0 ? prx::OP_APPLY : op == 1 ? prx::OP_DIR_APPLY : op == 2 ? prx::OP_CG : prx::OP_CG_FACES);
}

int prx_copy_bandwidth(size_t bytes, int reps, int device, double* ms) {
    if (bytes < 16 || reps <= 0 || !ms) return PRX_ERR_INVALID;
    try {
        using namespace prx;
        PRX_HIP(hipSetDevice(device));
        const int64_t n = (int64_t)(bytes / 2 / 16);
        float *a = nullptr, *b = nullptr;
        PRX_HIP(hipMalloc((void**)&a, n * 16));
        PRX_HIP(hipMalloc((void**)&b, n * 16));
        hipStream_t s;
        PRX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        PRX_HIP(hipMemsetAsync(a, 0, n * 16, s));
        hipEvent_t e0, e1;
        PRX_HIP(hipEventCreate(&e0));
        PRX_HIP(hipEventCreate(&e1));
        const int nb = (int)((n + 255) / 256);
        hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, s, a, b, n);
        PRX_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, s, a, b, n);
        PRX_HIP(hipEventRecord(e1, s));
        PRX_HIP(hipEventSynchronize(e1));
        float tms = 0.f;
        PRX_HIP(hipEventElapsedTime(&tms, e0, e1));
        *ms = (double)tms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        (void)hipFree(a);
        (void)hipFree(b);
        return PRX_OK;
    } catch (const prx::Error& e) {
        return e.code;
    }
}

int prx_brick_copy_bandwidth(int64_t nbricks, int streams, int precision, int reps, int device, double* ms) {
    if (nbricks < 8 || (nbricks & 7) || streams < 1 || streams > 3 || reps <= 0 || !ms) return PRX_ERR_INVALID;
    try {
        using namespace prx;
        PRX_HIP(hipSetDevice(device));
        const size_t esz = precision == PRX_PREC_FP64 ? 8 : 4;
        const size_t bytes = (size_t)nbricks * 4096 * esz;
        std::vector<void*> bufs;
        struct Free {
            std::vector<void*>& b;
            ~Free() { for (void* p : b) (void)hipFree(p); }
        } fr{bufs};
        for (int i = 0; i < 2 * streams; ++i) {
            void* q = nullptr;
            PRX_HIP(hipMalloc(&q, bytes));
            bufs.push_back(q);
        }
        hipStream_t s;
        PRX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        for (int i = 0; i < streams; ++i) PRX_HIP(hipMemsetAsync(bufs[i], 0, bytes, s));
        // the tile kernels' XCD chunk: consecutive bricks per XCD, at most 64
        uint32_t chunk = (uint32_t)(nbricks >> 3);
        while (chunk > 64 && (chunk & 1u) == 0) chunk >>= 1;
        hipEvent_t e0, e1;
        PRX_HIP(hipEventCreate(&e0));
        PRX_HIP(hipEventCreate(&e1));
        auto launch = [&]() {
            if (esz == 8) {
                BrickPtrs<double> p{};
                for (int i = 0; i < streams; ++i) { p.in[i] = (const double*)bufs[i]; p.out[i] = (double*)bufs[streams + i]; }
                if (streams == 1) hipLaunchKernelGGL((k_brick_copy<double, 1>), dim3((unsigned)nbricks), dim3(256), 0, s, p, chunk);
                else if (streams == 2) hipLaunchKernelGGL((k_brick_copy<double, 2>), dim3((unsigned)nbricks), dim3(256), 0, s, p, chunk);
                else hipLaunchKernelGGL((k_brick_copy<double, 3>), dim3((unsigned)nbricks), dim3(256), 0, s, p, chunk);
            } else {
                BrickPtrs<float> p{};
                for (int i = 0; i < streams; ++i) { p.in[i] = (const float*)bufs[i]; p.out[i] = (float*)bufs[streams + i]; }
                if (streams == 1) hipLaunchKernelGGL((k_brick_copy<float, 1>), dim3((unsigned)nbricks), dim3(256), 0, s, p, chunk);
                else if (streams == 2) hipLaunchKernelGGL((k_brick_copy<float, 2>), dim3((unsigned)nbricks), dim3(256), 0, s, p, chunk);
                else hipLaunchKernelGGL((k_brick_copy<float, 3>), dim3((unsigned)nbricks), dim3(256), 0, s, p, chunk);
            }
        };
        launch();
        PRX_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) launch();
        PRX_HIP(hipEventRecord(e1, s));
        PRX_HIP(hipEventSynchronize(e1));
        PRX_LAUNCH_CHECK();
        float tms = 0.f;
        PRX_HIP(hipEventElapsedTime(&tms, e0, e1));
        *ms = (double)tms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        return PRX_OK;
    } catch (const prx::Error& e) {
        return e.code;
    }
}

}  // extern "C"
