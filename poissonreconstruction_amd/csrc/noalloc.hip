// prx_csr_cg_noalloc_device: solve_cg (src/solvers.rs:20-69) on a CSR whose
// arrays, right-hand side and solution are already in device memory, with
// the solver's scratch (r, d, q, the partial sums and the CG scalars) in one
// fixed device array of this library instead of allocations.  It runs the
// same kernels in the same order as prx_csr_cg, so both give the same x and
// iteration count on the same input.
//
// The array `scratch` exists once per GPU for the life of the process.  Host
// code never passes `scratch` itself (in host code that name is a host-side
// stand-in, not a device address): the first call on a device, with that
// device current, looks up the device address with hipGetSymbolAddress and
// keeps it per device; kernels receive that address.  Every call on a device
// shares its array, so calls that overlap on one device (two host threads,
// or two streams) race and must be serialised by the caller.
#include <mutex>

#include "kernels.hpp"
#include "tree.hpp"

namespace {

constexpr int64_t kMaxRows = PRX_CSR_CG_NOALLOC_MAX_ROWS;
// partial sums: the reduction kernels leave at most 2048 (k_dot, k_sub) or
// n / 2048 (k_cg_update) partials for n <= kMaxRows
constexpr int64_t kPartials = 4096;
constexpr size_t kStateBytes = (sizeof(prx::CgState) + 255) / 256 * 256;
constexpr size_t kScratchBytes = (size_t)(3 * kMaxRows + kPartials) * sizeof(double) + kStateBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per device
void* scratch_on(int dev) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw prx::Error(PRX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

__global__ void k_state_init(prx::CgState* st, double eps, int64_t max_iters) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        prx::CgState h{};
        h.eps = eps;
        h.max_iters = max_iters;
        *st = h;
    }
}

__global__ void k_copy_f64(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

extern "C" {

int prx_csr_cg_noalloc_device(int64_t n, const int64_t* off, const int64_t* col, const double* val, const double* b,
                              double* x, uint64_t max_iters, double eps, uint64_t* iters, void* stream) {
    using namespace prx;
    PRX_API_BEGIN
    PRX_REQUIRE(n >= 0 && off && col && val && b && x, PRX_ERR_INVALID, "prx_csr_cg_noalloc_device: null pointer or negative size");
    PRX_REQUIRE(n <= kMaxRows, PRX_ERR_INVALID,
                "prx_csr_cg_noalloc_device: n = " + std::to_string(n) + " exceeds the fixed scratch of " +
                    std::to_string(kMaxRows) + " rows (PRX_CSR_CG_NOALLOC_MAX_ROWS)");
    int dev = -1;
    PRX_HIP(hipGetDevice(&dev));
    PRX_REQUIRE(dev >= 0 && dev < kMaxDevices, PRX_ERR_DEVICE, "device ordinal out of range");
    unsigned char* base = (unsigned char*)scratch_on(dev);
    double* dr = (double*)base;
    double* dd = dr + kMaxRows;
    double* dq = dd + kMaxRows;
    double* part = dq + kMaxRows;
    CgState* st = (CgState*)(part + kPartials);
    hipStream_t s = (hipStream_t)stream;

    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st, eps, (int64_t)max_iters);
    PRX_LAUNCH_CHECK();
    int nb = 0;
    launch_csr_spmv(n, off, col, val, x, dq, s);
    launch_sub<double>(dr, b, dq, n, part, &nb, nullptr, s);
    launch_fin(st, part, nb, FIN_INIT_REF, 0, s);
    hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536))),
                       dim3(256), 0, s, dd, dr, n);
    PRX_LAUNCH_CHECK();
    CgState h{};
    for (uint64_t it = 0; it < max_iters; ++it) {
        launch_csr_spmv(n, off, col, val, dd, dq, s);
        launch_dot<double>(dd, dq, n, part, &nb, s);
        launch_fin(st, part, nb, FIN_ALPHA, 0, s);
        if ((it + 1) % 10 == 0) {
            launch_axpy_alpha<double>(x, nullptr, dd, st, n, s);
            launch_csr_spmv(n, off, col, val, x, dq, s);
            launch_sub<double>(dr, b, dq, n, part, &nb, st, s);
        } else {
            launch_cg_update<double>(x, dr, nullptr, dd, dq, st, n, part, &nb, s);
        }
        launch_fin(st, part, nb, FIN_BETA, 0, s);
        launch_xpby_beta<double>(dd, dr, st, n, s);
        if ((it & 15) == 15) {
            PRX_HIP(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, s));
            PRX_HIP(hipStreamSynchronize(s));
            if (h.done) break;
        }
    }
    PRX_HIP(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, s));
    PRX_HIP(hipStreamSynchronize(s));
    if (iters) *iters = h.done ? (uint64_t)h.iters : max_iters;
    PRX_API_END
}

}  // extern "C"
