// Internal declarations shared by the HIP translation units of libprx.
//
// Grid layout: the reference's complete octree numbers its depth-D leaves in
// DFS order, which is the Morton code of the leaf offset with x in bit 0
// (src/octree.rs:145-153, :235-248; src/fem_tree.rs:104-121).  Every fine
// vector here is stored in that order, so the index of node (x,y,z) is
// morton3(x,y,z) and an aligned B^3 block of nodes is a contiguous range.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/prx.h"

namespace prx {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PRX_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::prx::Error(e_ == hipErrorOutOfMemory ? PRX_ERR_OOM : PRX_ERR_DEVICE,       \
                               std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

#define PRX_LAUNCH_CHECK() PRX_HIP(hipGetLastError())

// ------------------------------------------------------------------ Morton
__host__ __device__ inline uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__host__ __device__ inline uint32_t compact3(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0x030000FFu;
    v = (v | (v >> 16)) & 0x3ffu;
    return v;
}
__host__ __device__ inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
__host__ __device__ inline void demorton3(uint32_t m, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = compact3(m);
    y = compact3(m >> 1);
    z = compact3(m >> 2);
}

// Rust `f as isize` for the ranges the path produces: saturating, NaN -> 0.
__host__ __device__ inline int64_t sat_i64(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}
__host__ __device__ inline int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }

// ------------------------------------------------------------------ stencil
// Constant coefficients of one depth: K(0)=centre, K(edge), K(corner)
// (face entries are exactly 0.0 and dropped, src/fem_tree.rs:189-195),
// the global screening mass wg on the diagonal (src/fem_tree.rs:203-204) and
// wh3 for the per-point blocks.  For the FP32 operator all of them are
// divided by h, which keeps the system O(1) without changing its solution.
template <typename T>
struct Coef {
    T c0, ce, cc, wg, wh3;
};

// 16-byte vector of T (one global load/store instruction per lane).
template <typename T> struct V16;
template <> struct V16<float> { using type = float4; static constexpr int n = 4; };
template <> struct V16<double> { using type = double2; static constexpr int n = 2; };

// Native 16-byte vectors for the non-temporal load/store builtins: streamed
// data (read once, written once per pass) bypasses the cache hierarchy's
// normal allocation, which measurably raises streaming bandwidth on gfx950.
template <typename T> struct NV16;
template <> struct NV16<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct NV16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename T>
__device__ inline typename NV16<T>::type ld_nt(const T* p) {
    return __builtin_nontemporal_load((const typename NV16<T>::type*)p);
}
template <typename T>
__device__ inline void st_nt(T* p, typename NV16<T>::type v) {
    __builtin_nontemporal_store(v, (typename NV16<T>::type*)p);
}
template <typename T>
__device__ inline typename NV16<T>::type ld_v(const T* p) {
    return *(const typename NV16<T>::type*)p;
}

// Touched-node lists: one entry per (point, corner) that reaches a node,
// sorted by node and, within a node, by input point order.
// CG direction d_new = r + beta*d_old as every kernel forms it (the same
// expression wherever a node value is formed, so a node read as halo by a
// neighbouring workgroup is bit-identical to its own interior value): one
// FP64 fma, rounded once to the vector type (an FP32 fma with beta rounded
// to FP32 measurably costs CG iterations of the FP32 correction system).
__device__ __forceinline__ double dir_form(double bt, double d, double r) { return fma(bt, d, r); }
__device__ __forceinline__ float dir_form(double bt, float d, float r) { return (float)fma(bt, (double)d, (double)r); }
// The CG residual update r - a q, in FP64 and rounded once, formed the same
// way wherever a node's updated residual is needed.
__device__ __forceinline__ double r_form(double a, double q, double r) { return fma(-a, q, r); }
__device__ __forceinline__ float r_form(double a, float q, float r) { return (float)fma(-a, (double)q, (double)r); }
// OP_CG operand: d' = r' + bt d with r' = r - a q when rp (else r' = r)
template <typename T>
__device__ __forceinline__ T cg_form(double bt, double a, int rp, T d, T r, T q) {
    return dir_form(bt, d, rp ? r_form(a, q, r) : r);
}

struct NodeList {
    int mode = 0;                 // 0 = floor (splat), 1 = trunc (screening)
    int64_t T = 0;                // touched nodes
    int64_t M = 0;                // entries
    uint32_t* node = nullptr;     // [T] node index (Morton)
    uint32_t* off = nullptr;      // [T+1] entry offsets
    uint32_t* code = nullptr;     // [M] 8*p + corner, p = input point index
    uint32_t* sp = nullptr;       // [M] sorted position of the point (screening lookups)
    uint32_t* spc = nullptr;      // [M] 8*sp + corner (the weight is recomputed from the point)
    double* phi = nullptr;        // [M] trilinear weight (bx*by)*bz
    float* phi32 = nullptr;       // [M]
    double* diag = nullptr;       // [T] sum of phi^2 (screening diagonal without wh3)
    uint32_t* brick = nullptr;    // [nbricks+1] first touched node of each 16^3 brick
    int brick_log = 12;           // log2 of nodes per brick
    uint32_t* brick8 = nullptr;   // [N/512+1] the same for 8^3 bricks (D >= 3)
};

// The occupied trunc cells of the sorted points and, per touched node of the
// screening list, its (cell, corner) pairs: the face-form CG pass forms the
// screening term cell by cell (every point of a cell has the same 8 corners).
struct CellLists {
    int64_t C = 0;              // occupied cells
    uint32_t* cstart = nullptr; // [C+1] first sorted point of each cell
    uint32_t* coff = nullptr;   // [T+1] per touched node: its range in cval
    uint32_t* cval = nullptr;   // [8C] 8 * cell + corner, grouped by node (cell order inside a node)
};


}  // namespace prx
