// Internal helpers shared by the neuwon_amd HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/neuwon_amd.h"

namespace nw {

// Thread-local message of the last failure (returned by nw_last_error).
void set_error(const char* fmt, ...);

#define NW_HIP(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::nw::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),         \
                            __FILE__, __LINE__);                                           \
            return NW_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

#define NW_REQUIRE(cond, ...)                                                              \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            ::nw::set_error(__VA_ARGS__);                                                  \
            return NW_ERR_ARG;                                                             \
        }                                                                                  \
    } while (0)

constexpr int kWave = 64;

// Voltages are held in LDS as FP32 offsets from this reference (mV): the offset keeps the FP32
// rounding proportional to the distance from rest instead of to the full membrane potential.
constexpr double kVRef = -65.0;

// Sum across the 64 lanes of a wavefront; every lane receives the total.
__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    return x;
}
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    return x;
}

// v = P x for one CSR row held by one wavefront: x is stored in LDS as FP32 offsets from kVRef,
// products of the FP32 values and offsets are accumulated in FP64, and the reference part is
// restored with the exact FP64 sum of the stored row:  P x = kVRef * rowsum + P (x - kVRef).
__device__ __forceinline__ double row_apply(const uint32_t* __restrict__ roff,
                                            const float* __restrict__ vals,
                                            const uint16_t* __restrict__ cols,
                                            const float* __restrict__ xoff, double rowsum, int r,
                                            int lane) {
    const uint32_t beg = roff[r], end = roff[r + 1];
    double acc0 = 0.0, acc1 = 0.0;
    uint32_t k = beg + lane;
    for (; k + kWave < end; k += 2 * kWave) {
        const float a = vals[k], b = vals[k + kWave];
        const uint16_t ca = cols[k], cb = cols[k + kWave];
        acc0 = fma((double)a, (double)xoff[ca], acc0);
        acc1 = fma((double)b, (double)xoff[cb], acc1);
    }
    if (k < end) acc0 = fma((double)vals[k], (double)xoff[cols[k]], acc0);
    return fma(kVRef, rowsum, wave_sum(acc0 + acc1));
}

// Location of `value` in a linear input table, reference semantics of the generated LTI kernel
// (lti_sim/codegen.py:44-59): u = (value - vmin) * nb / (vmax - vmin); b = trunc(u);
// b > nb-1 -> (nb-1, 1); b < 0 -> (0, 0); otherwise frac = u - b.
template <typename T>
__device__ __forceinline__ void bucket_of(T value, T vmin, T bucket_frq, int nb, int& b, T& frac) {
    T u = (value - vmin) * bucket_frq;
    int bi = (int)u;  // truncation toward zero, as int() in the reference
    if (bi > nb - 1) {
        b = nb - 1;
        frac = T(1);
    } else if (bi < 0) {
        b = 0;
        frac = T(0);
    } else {
        b = bi;
        frac = u - (T)bi;
    }
}

// Counter-based generator of Poisson pulse onsets (seed, cell, tick) -> uint32.
__host__ __device__ __forceinline__ uint32_t stim_hash(uint32_t seed, uint32_t cell, uint32_t tick) {
    uint32_t h = (seed * 0x9E3779B1u) ^ (cell * 0x85EBCA77u) ^ (tick * 0xC2B2AE3Du);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

}  // namespace nw

// Device-side layout of a set of per-cell CSR propagators.
struct nw_propagator {
    int device = 0;
    int64_t num_cells = 0;
    int64_t num_rows = 0;
    int64_t nnz = 0;
    int32_t max_cell_rows = 0;
    int32_t max_row_nnz = 0;
    double build_seconds = 0.0;
    // device
    int64_t* cell_row0 = nullptr;  // [num_cells+1] first global row of each cell
    int64_t* cell_nz0 = nullptr;   // [num_cells]   first stored entry of each cell
    uint32_t* row_off = nullptr;   // cell c: row_off[cell_row0[c] + c + r], r = 0..n_c (local)
    float* vals = nullptr;         // [nnz]
    uint16_t* cols = nullptr;      // [nnz] column local to the cell
    double* rowsum = nullptr;      // [num_rows] exact FP64 sum of every stored FP32 row
    int32_t* order = nullptr;      // cells by decreasing stored entries (launch order)
    // host mirrors
    std::vector<int64_t> h_cell_row0, h_cell_nz0;
    int64_t vals_capacity = 0;
};
