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

// Voltages are held in LDS as FP32 offsets from a per-cell reference voltage (the first segment's
// value when the launch starts): the FP32 rounding is then proportional to the spread of voltages
// within the cell instead of to the membrane potential itself.

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

// Stored rows are padded to a multiple of kRowAlign entries (padding: value 0, column = the row),
// so a lane can fetch kRowAlign consecutive values (16 B) and columns (8 B) with one vector load.
constexpr int kRowAlign = 4;
constexpr int kRowUnroll = 2;  // vector loads in flight per lane per round (512 entries per wave)

struct RowChunk {
    float4 a[kRowUnroll];
    uint2 c[kRowUnroll];  // four u16 columns
};

__device__ __forceinline__ void row_load(RowChunk& ch, const float* __restrict__ vals,
                                         const uint16_t* __restrict__ cols, uint32_t base,
                                         uint32_t end, int lane) {
#pragma unroll
    for (int u = 0; u < kRowUnroll; ++u) {
        const uint32_t k = base + (uint32_t)(u * kWave + lane) * kRowAlign;
        if (k < end) {
            ch.a[u] = *reinterpret_cast<const float4*>(vals + k);
            ch.c[u] = *reinterpret_cast<const uint2*>(cols + k);
        } else {
            ch.a[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            ch.c[u] = make_uint2(0u, 0u);
        }
    }
}

__device__ __forceinline__ double row_consume(const RowChunk& ch, const float* __restrict__ xoff,
                                              double acc) {
#pragma unroll
    for (int u = 0; u < kRowUnroll; ++u) {
        const uint32_t c01 = ch.c[u].x, c23 = ch.c[u].y;
        acc = fma((double)ch.a[u].x, (double)xoff[c01 & 0xFFFFu], acc);
        acc = fma((double)ch.a[u].y, (double)xoff[c01 >> 16], acc);
        acc = fma((double)ch.a[u].z, (double)xoff[c23 & 0xFFFFu], acc);
        acc = fma((double)ch.a[u].w, (double)xoff[c23 >> 16], acc);
    }
    return acc;
}

// out[r] = (P x)_r - vref for the rows r = first, first + stride, ... of one cell, one wavefront
// per row.  x is stored in LDS as FP32 offsets from vref, products accumulate in FP64 and the
// reference part is restored with the exact FP64 sum of the stored row:
//     P x = vref * rowsum + P (x - vref).
// The next row's offsets and first chunk are fetched before the current row is reduced.
__device__ __forceinline__ void spmv_rows(const uint32_t* __restrict__ roff,
                                          const float* __restrict__ vals,
                                          const uint16_t* __restrict__ cols,
                                          const double* __restrict__ rowsum,
                                          const float* __restrict__ xoff, float* __restrict__ out,
                                          double vref, int nrows, int first, int stride, int lane) {
    int r = first;
    if (r >= nrows) return;
    uint32_t beg = roff[r], end = roff[r + 1];
    RowChunk ch;
    row_load(ch, vals, cols, beg, end, lane);
    constexpr uint32_t kChunk = kWave * kRowAlign * kRowUnroll;
    while (true) {
        const int rn = r + stride;
        uint32_t nbeg = 0, nend = 0;
        if (rn < nrows) {
            nbeg = roff[rn];
            nend = roff[rn + 1];
        }
        double acc = row_consume(ch, xoff, 0.0);
        for (uint32_t base = beg + kChunk; base < end; base += kChunk) {
            row_load(ch, vals, cols, base, end, lane);
            acc = row_consume(ch, xoff, acc);
        }
        if (rn < nrows) row_load(ch, vals, cols, nbeg, nend, lane);
        const double s = fma(vref, rowsum[r], wave_sum(acc));
        if (lane == 0) out[r] = (float)(s - vref);
        if (rn >= nrows) break;
        r = rn;
        beg = nbeg;
        end = nend;
    }
}

// Same product with one thread per row, for propagators whose rows hold a few entries (the
// diffusion propagator): rows are padded to kRowAlign, so every load is a 16-B / 8-B vector load.
__device__ __forceinline__ void spmv_rows_lane(const uint32_t* __restrict__ roff,
                                               const float* __restrict__ vals,
                                               const uint16_t* __restrict__ cols,
                                               const double* __restrict__ rowsum,
                                               const float* __restrict__ xoff,
                                               float* __restrict__ out, double vref, int nrows,
                                               int tid, int nthreads) {
    for (int r = tid; r < nrows; r += nthreads) {
        const uint32_t beg = roff[r], end = roff[r + 1];
        double acc = 0.0;
        for (uint32_t k = beg; k < end; k += kRowAlign) {
            const float4 a = *reinterpret_cast<const float4*>(vals + k);
            const uint2 c = *reinterpret_cast<const uint2*>(cols + k);
            acc = fma((double)a.x, (double)xoff[c.x & 0xFFFFu], acc);
            acc = fma((double)a.y, (double)xoff[c.x >> 16], acc);
            acc = fma((double)a.z, (double)xoff[c.y & 0xFFFFu], acc);
            acc = fma((double)a.w, (double)xoff[c.y >> 16], acc);
        }
        out[r] = (float)(fma(vref, rowsum[r], acc) - vref);
    }
}

// Rows of at most this many stored entries use one thread per row.
constexpr int kLaneRowMax = 32;

template <int ROWMODE>
__device__ __forceinline__ void spmv(const uint32_t* __restrict__ roff,
                                     const float* __restrict__ vals,
                                     const uint16_t* __restrict__ cols,
                                     const double* __restrict__ rowsum,
                                     const float* __restrict__ xoff, float* __restrict__ out,
                                     double vref, int nrows) {
    if (ROWMODE == 1) {
        spmv_rows_lane(roff, vals, cols, rowsum, xoff, out, vref, nrows, threadIdx.x, blockDim.x);
    } else {
        const int lane = threadIdx.x & (kWave - 1);
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        spmv_rows(roff, vals, cols, rowsum, xoff, out, vref, nrows, wave, blockDim.x / kWave, lane);
    }
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
    int64_t nnz = 0;               // entries kept by the pruning
    int64_t stored = 0;            // entries stored, rows padded to kRowAlign
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
