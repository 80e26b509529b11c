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
    if (!t || reps <= 0 || op < 0 || op > 2) return PRX_ERR_INVALID;
    return prx::time_operator(t, precision, reps, ms_gather, ms_stencil,
                              op == 0 ? prx::OP_APPLY : op == 1 ? prx::OP_DIR_APPLY : prx::OP_CG);
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

}  // extern "C"
