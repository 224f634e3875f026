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

// Sum across the 64 lanes with DPP row shifts and row broadcasts (register moves, no LDS round
// trip): within every row of 16 lanes shifts by 1, 2, 4, 8 leave the row's sum in its lane 15,
// row_bcast:15 and row_bcast:31 carry those into lane 63.  Returns the total in every lane.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double x) {
    const int2 b = __builtin_bit_cast(int2, x);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(0, b.x, CTRL, ROW_MASK, 0xf, true);
    r.y = __builtin_amdgcn_update_dpp(0, b.y, CTRL, ROW_MASK, 0xf, true);
    return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ double wave_total(double x) {
    x += dpp_f64<0x111, 0xf>(x);  // row_shr:1
    x += dpp_f64<0x112, 0xf>(x);  // row_shr:2
    x += dpp_f64<0x114, 0xf>(x);  // row_shr:4
    x += dpp_f64<0x118, 0xf>(x);  // row_shr:8
    x += dpp_f64<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
    x += dpp_f64<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
    const int2 b = __builtin_bit_cast(int2, x);
    int2 t;
    t.x = __builtin_amdgcn_readlane(b.x, 63);
    t.y = __builtin_amdgcn_readlane(b.y, 63);
    return __builtin_bit_cast(double, t);
}

// Stored rows are padded to a multiple of kRowAlign entries (padding: value 0, column = the row),
// so a lane can fetch kRowAlign consecutive values (16 B) and columns (8 B) with one vector load.
constexpr int kRowAlign = 4;
// Vector loads per lane per chunk: 2 (512 entries per wave) for the general format, whose rows hold
// ~540 entries at the 1e-9 pruning; 1 (256) for the symmetric format's ~290-entry rows.
constexpr int kRowUnrollGeneral = 2;
constexpr int kRowUnrollSym = 1;
// Chunk buffers per wave (chunks in flight + 1).
constexpr int kBuffersGeneral = 2;
constexpr int kBuffersSym = 3;

// Order of a row's entries in memory.  Lane l of a wavefront loads the 4 entries at 4l..4l+3 of
// each 256-entry chunk; the build stores the chunk's k-th entry (column order) at slot
//     4 (k mod m) + k / m,   m = (entries in the chunk) / 4,
// so that for every q the lanes' q-th entries are consecutive columns: the LDS gathers of x (and
// the scatters of the symmetric product) then touch consecutive banks instead of every fourth.
__host__ __device__ __forceinline__ uint32_t stored_slot(uint32_t k, uint32_t row_len) {
    constexpr uint32_t kChunkEntries = 64u * kRowAlign;
    const uint32_t c0 = k - k % kChunkEntries;
    const uint32_t len = (row_len - c0 < kChunkEntries) ? row_len - c0 : kChunkEntries;
    const uint32_t m = len / kRowAlign;
    const uint32_t i = k - c0;
    return c0 + kRowAlign * (i % m) + i / m;
}

template <int U>
struct RowChunk {
    float4 a[U];
    uint2 c[U];  // four u16 columns
};

// Loads one chunk at `base` of a row ending at `end`.  Every lane loads: lanes past the end of the
// row reload the chunk's first entry (in bounds, every stored row has at least kRowAlign entries)
// and row_valid() masks them out when the chunk is consumed.  Loads issued unconditionally let
// the compiler wait for exactly one chunk while the next is in flight.
template <int U>
__device__ __forceinline__ void row_load(RowChunk<U>& ch, const float* __restrict__ vals,
                                         const uint16_t* __restrict__ cols, uint32_t base,
                                         uint32_t end, int lane) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t k = base + (uint32_t)(u * kWave + lane) * kRowAlign;
        const uint32_t kk = (k < end) ? k : base;
        ch.a[u] = *reinterpret_cast<const float4*>(vals + kk);
        ch.c[u] = *reinterpret_cast<const uint2*>(cols + kk);
    }
}

// Whether part u of a lane's chunk lies inside the row (`remaining` = end - base of the chunk).
__device__ __forceinline__ bool row_valid(int u, int lane, uint32_t remaining) {
    return (uint32_t)(u * kWave + lane) * kRowAlign < remaining;
}

template <int U>
__device__ __forceinline__ double row_consume(const RowChunk<U>& ch, const float* __restrict__ xoff,
                                              uint32_t remaining, int lane, double acc) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const bool ok = row_valid(u, lane, remaining);
        const float4 a = ok ? ch.a[u] : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t c01 = ch.c[u].x, c23 = ch.c[u].y;
        acc = fma((double)a.x, (double)xoff[c01 & 0xFFFFu], acc);
        acc = fma((double)a.y, (double)xoff[c01 >> 16], acc);
        acc = fma((double)a.z, (double)xoff[c23 & 0xFFFFu], acc);
        acc = fma((double)a.w, (double)xoff[c23 >> 16], acc);
    }
    return acc;
}

// Walks the stored rows first, first + stride, ... of one cell, one wavefront per row, in chunks of
// kChunk entries, with NB - 1 chunks (NB = 2..4 buffers) always loading while one is consumed.  The
// buffers are used in turn, so the loads' destination registers stay fixed and consuming one
// chunk waits only for its own loads; every step issues its load (past the last chunk the first
// one is loaded again, unused), which keeps those waits exact.  The offsets of the next row to
// enter the pipeline (and rowval of it, with HAS_VAL) are fetched one row ahead.
//   consume(chunk, r, remaining, acc) -> acc   products of one chunk of row r
//   finish(r, acc, rowval[r])                  after the row's last chunk (acc: this lane's sum)
struct ChunkPos {
    int r;          // row; >= nrows once every row is done
    uint32_t base;  // first entry of the chunk
    uint32_t end;   // end of the row
    double val;     // rowval[r]
};

template <int U, int NB, bool HAS_VAL, typename Consume, typename Finish>
__device__ __forceinline__ void stream_rows(const uint32_t* __restrict__ roff,
                                            const float* __restrict__ vals,
                                            const uint16_t* __restrict__ cols,
                                            const double* __restrict__ rowval, int nrows,
                                            int first, int stride, int lane, Consume consume,
                                            Finish finish) {
    static_assert(NB >= 2 && NB <= 4, "two to four chunk buffers");
    constexpr uint32_t kChunk = kWave * kRowAlign * U;
    if (first >= nrows) return;
    // row offsets and values are read-only for the kernel's lifetime and indexed by the
    // wave-uniform row: read them through the constant address space (scalar loads, which do not
    // enter the vector-memory counter that the chunk loads are waited on with)
    using ConstU32 = const __attribute__((address_space(4))) uint32_t*;
    using ConstF64 = const __attribute__((address_space(4))) double*;
    const ConstU32 ro = (ConstU32)roff;
    const ConstF64 rv = (ConstF64)rowval;
    auto row_pos = [&](int r) -> ChunkPos {  // first chunk of row r (loads clamped in bounds)
        const int rc = min(r, nrows - 1);
        return ChunkPos{r, ro[rc], ro[rc + 1], HAS_VAL ? (double)rv[rc] : 0.0};
    };
    const ChunkPos p0 = row_pos(first);
    ChunkPos pre = row_pos(first + stride);
    auto next = [&](const ChunkPos& p) -> ChunkPos {
        if (p.r >= nrows) return p;
        if (p.base + kChunk < p.end) return ChunkPos{p.r, p.base + kChunk, p.end, p.val};
        const ChunkPos q = pre;
        pre = row_pos(min(q.r + stride, nrows));
        return q;
    };
    auto load = [&](RowChunk<U>& ch, const ChunkPos& p) {
        const bool live = p.r < nrows;
        row_load(ch, vals, cols, live ? p.base : p0.base, live ? p.end : p0.end, lane);
    };
    RowChunk<U> buf[NB];
    ChunkPos d[NB - 1];  // d[0]: the chunk consumed next; d[1..]: chunks already loading
    d[0] = p0;
    load(buf[0], d[0]);
#pragma unroll
    for (int i = 1; i < NB - 1; ++i) {
        d[i] = next(d[i - 1]);
        load(buf[i], d[i]);
    }
    double acc = 0.0;
    // consume `cur` (chunk d[0]) after issuing the load of the chunk NB - 1 ahead into `ld`
    auto step = [&](RowChunk<U>& cur, RowChunk<U>& ld) -> bool {
        const ChunkPos dn = next(d[NB - 2]);
        load(ld, dn);
        acc = consume(cur, d[0].r, d[0].end - d[0].base, acc);
        if (d[0].base + kChunk >= d[0].end) {
            finish(d[0].r, acc, d[0].val);
            acc = 0.0;
        }
#pragma unroll
        for (int i = 0; i < NB - 2; ++i) d[i] = d[i + 1];
        d[NB - 2] = dn;
        return d[0].r < nrows;
    };
    if constexpr (NB == 2) {
        while (step(buf[0], buf[1]) && step(buf[1], buf[0])) {
        }
    } else if constexpr (NB == 3) {
        while (step(buf[0], buf[2]) && step(buf[1], buf[0]) && step(buf[2], buf[1])) {
        }
    } else {
        while (step(buf[0], buf[3]) && step(buf[1], buf[0]) && step(buf[2], buf[1]) &&
               step(buf[3], buf[2])) {
        }
    }
}

// out[r] = (P x)_r - vref for the rows r = first, first + stride, ... of one cell, one wavefront
// per row.  x is stored in LDS as FP32 offsets from vref, products accumulate in FP64 and the
// reference part is restored with the exact FP64 sum of the stored row:
//     P x = vref * rowsum + P (x - vref).
__device__ __forceinline__ void spmv_rows(const uint32_t* __restrict__ roff,
                                          const float* __restrict__ vals,
                                          const uint16_t* __restrict__ cols,
                                          const double* __restrict__ rowsum,
                                          const float* __restrict__ xoff, float* __restrict__ out,
                                          double vref, int nrows, int first, int stride, int lane) {
    stream_rows<kRowUnrollGeneral, kBuffersGeneral, true>(
        roff, vals, cols, rowsum, nrows, first, stride, lane,
        [&](const RowChunk<kRowUnrollGeneral>& ch, int, uint32_t remaining, double acc) {
            return row_consume(ch, xoff, remaining, lane, acc);
        },
        [&](int r, double acc, double rs) {
            const double s = fma(vref, rs, wave_total(acc));
            if (lane == 0) out[r] = (float)(s - vref);
        });
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

// Symmetric format: z += S y with S stored on and above the diagonal.  One wavefront per stored
// row: the row part is reduced across the lanes, the transposed part (S_rj y_r into z_j, j > r)
// goes through LDS FP64 atomics.  z must be zeroed before and is complete after the block's
// barrier.
__device__ __forceinline__ void spmv_sym_rows(const uint32_t* __restrict__ roff,
                                              const float* __restrict__ vals,
                                              const uint16_t* __restrict__ cols,
                                              const float* __restrict__ ys, double* zs, int nrows,
                                              int first, int stride, int lane) {
    stream_rows<kRowUnrollSym, kBuffersSym, false>(
        roff, vals, cols, nullptr, nrows, first, stride, lane,
        [&](const RowChunk<kRowUnrollSym>& ch, int r, uint32_t remaining, double acc) {
            const double yr = ys[r];
#pragma unroll
            for (int u = 0; u < kRowUnrollSym; ++u) {
                const bool ok = row_valid(u, lane, remaining);
                const uint32_t cc[4] = {ch.c[u].x & 0xFFFFu, ch.c[u].x >> 16, ch.c[u].y & 0xFFFFu,
                                        ch.c[u].y >> 16};
                const float av[4] = {ch.a[u].x, ch.a[u].y, ch.a[u].z, ch.a[u].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double sv = ok ? (double)av[q] : 0.0;
                    acc = fma(sv, (double)ys[cc[q]], acc);
                    if (ok && cc[q] != (uint32_t)r) atomicAdd(&zs[cc[q]], sv * yr);
                }
            }
            return acc;
        },
        [&](int r, double acc, double) {
            acc = wave_total(acc);
            if (lane == 0) atomicAdd(&zs[r], acc);
        });
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
    // symmetric format: rows hold S = W P W^-1 (W = diag(w), w = sqrt(C)) on and above the
    // diagonal; P x = W^-1 S W x
    int symmetric = 0;
    double* wgt = nullptr;         // [num_rows] w of every segment (symmetric format)
    int32_t* order = nullptr;      // cells by decreasing stored entries (launch order)
    // Chebyshev store: neighbours of every row (local index, coefficient (2/rho) b_ij), slot 2
    // holding -(start + 1) of an overflow list ended by index -1 when a row has more than 3
    int chebyshev = 0;
    int32_t* nb_idx = nullptr;     // [num_rows][3]
    float* nb_coef = nullptr;      // [num_rows][3]
    int32_t* ov_idx = nullptr;     // overflow lists
    float* ov_coef = nullptr;
    int32_t* cheb_K = nullptr;     // [num_cells] index of the last series coefficient
    int64_t* cheb_off = nullptr;   // [num_cells] first coefficient of the cell
    double* cheb_c = nullptr;      // coefficients of all cells
    double* cheb_v64 = nullptr;    // [num_rows] voltages between the launches of a step
    // Cells whose rows have at most 3 neighbours (every binary tree) also run in depth-first
    // preorder, a first child right after its parent, laid out in slots: thread t of the workgroup
    // owns slots J t .. J t + J - 1 (J = ceil(slots / 1024)), slot s = J t + j lives at LDS place
    // j * 1024 + t.  A few slots stay empty (holes) so no row needs more than two LDS reads per
    // term.  pre_pos[i]: LDS place of natural row i.  Per LDS place of the cell (from pre_base[c]):
    // pre_far = places of the row's two far neighbours (| second << 16; kChebHole for an empty
    // slot), pre_coef = (2/rho) b of the neighbour in slot s - 1, in slot s + 1, and of the two far
    // ones.  pre_base[c] = -1: the cell runs in natural order.
    uint16_t* pre_pos = nullptr;   // [num_rows]
    uint32_t* pre_far = nullptr;   // [sum of 1024 J over preorder cells]
    float4* pre_coef = nullptr;    // [same]
    int64_t* pre_base = nullptr;   // [num_cells]
    int32_t* pre_size = nullptr;   // [num_cells] slots in use (rows and holes) | J << 16
    int64_t pre_cells = 0;         // cells laid out in preorder
    int64_t pre_places = 0;        // entries of pre_far / pre_coef
    int32_t cheb_maxj = 0;         // row groups per thread the store's cells need (1..kChebRows)
    int32_t cheb_small = 0;        // every cell preorder, one row group of <= kChebSmallBlock slots
    // Rational store (also chebyshev = 1: its launches, LDS places and pre_pos / pre_size): rows
    // of a cell in level order (root, its children, theirs, ...; a parent's children adjacent),
    // pre_pos[i] = level place of natural row i, pre_base[c] = 0 for every cell.
    int rational = 0;
    uint32_t* rat_link = nullptr;  // [num_rows, level order] parent | first child << 13 | children << 26
    float2* rat_b = nullptr;       // [num_rows, level order] (B[parent][row], B[row][parent])
    double2* rat_invd = nullptr;   // [kRatPoles][num_rows, level order] 1 / pivot of (B - theta I)
    uint16_t* rat_lev = nullptr;   // level starts of every cell (levels + 1 each)
    int64_t* rat_lev_off = nullptr;  // [num_cells] first level start of the cell
    int32_t* rat_depth = nullptr;  // [num_cells] levels
    int32_t rat_max_levels = 0;
    double rat_mean_levels = 0.0;
    double mean_terms = 0.0;
    // Graphs of nw_hh_step's half-step launches on the matrix-free stores (NW_LAUNCH_GRAPH): one
    // per argument set, least recently used dropped beyond kStepGraphs; the kernels of a graph
    // read the first step's index from step_dev, written before every replay.
    int64_t* step_dev = nullptr;   // [1]
    struct StepGraph {
        std::vector<unsigned char> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    mutable std::vector<StepGraph> graphs;
    mutable uint64_t graph_clock = 0;
    mutable hipStream_t capture_stream = nullptr;
    // host mirrors
    std::vector<int64_t> h_cell_row0, h_cell_nz0;
    std::vector<int32_t> h_cheb_K;
    std::vector<int64_t> h_cheb_off;
    std::vector<double> h_cheb_c, h_rho;
    int64_t vals_capacity = 0;
};

constexpr int kStepGraphs = 4;

// Device pointers of a Chebyshev store, passed to the kernels that read it.
struct ChebStore {
    const int32_t* nb_idx;
    const float* nb_coef;
    const int32_t* ov_idx;
    const float* ov_coef;
    const int32_t* K;
    const int64_t* off;
    const double* c;
    const uint16_t* pre_pos;
    const uint32_t* pre_far;
    const float4* pre_coef;
    const int64_t* pre_base;
    const int32_t* pre_size;
    int32_t rational;
    int64_t num_rows;
    const uint32_t* rat_link;
    const float2* rat_b;
    const double2* rat_invd;
    const uint16_t* rat_lev;
    const int64_t* rat_lev_off;
    const int32_t* rat_depth;
};

inline ChebStore cheb_store(const nw_propagator* p) {
    return ChebStore{p->nb_idx,   p->nb_coef,  p->ov_idx,    p->ov_coef,     p->cheb_K,
                     p->cheb_off, p->cheb_c,   p->pre_pos,   p->pre_far,     p->pre_coef,
                     p->pre_base, p->pre_size, p->rational,  p->num_rows,    p->rat_link,
                     p->rat_b,    p->rat_invd, p->rat_lev,   p->rat_lev_off, p->rat_depth};
}

// The rational store's approximation exp(x) ~ a0 + 2 Re sum_j alpha_j / (x - theta_j) on
// (-inf, 0]: type (10, 10), poles from the AAA algorithm on exp(9 (t - 1) / (t + 1)), t in
// [-1, 1), residues by least squares, a0 moved so that r(0) = 1; maximum error 2.9e-10.
// Layout: a0, then theta_re, theta_im, alpha_re, alpha_im per pole.
constexpr int kRatPoles = 5;
#define NW_RAT_COEFFICIENTS                                                                    \
    2.802151606726937e-11,                                                                     \
    -5.056635951881228, 11.421609290485412, -0.006437906767520651, 0.0005054310548602677,      \
    -0.8025731342599113, 8.606156657901737, 0.2987978122522727, 0.02160797576074506,           \
    1.803811562314832, 6.052904140355052, -2.7893832883191694, -1.363533660934761,             \
    3.369248005747136, 3.602670710938092, 7.700048093610149, 9.67073157332556,                 \
    4.11129462003477, 1.1966884849799233, -5.203024560624244, -22.951290497759175
constexpr double kRatHost[1 + 4 * kRatPoles] = {NW_RAT_COEFFICIENTS};
static __constant__ double kRatDev[1 + 4 * kRatPoles] = {NW_RAT_COEFFICIENTS};

// Rows per thread of the Chebyshev products (cells up to kChebBlock * kChebRows segments).
constexpr int kChebBlock = 1024;
constexpr int kChebRows = 5;
// Workgroup of the launches over stores of small cells (every cell preorder, one row group, at
// most this many slots): four workgroups per CU instead of two 1024-thread ones half idle.
constexpr int kChebSmallBlock = 512;

// One Chebyshev product v = sum_k c_k T_k(A~) x for the rows this thread owns (row
// threadIdx.x + j * kChebBlock), A~ t = t + sum_j coef_ij (t_j - t_i) over the row's neighbours.
// On entry buf0[i] = x_i for every row of the cell (after a barrier); on return the thread's
// rows hold their result in `out[j]` and the call has ended with a barrier, so both LDS buffers
// are free.  K = 0 returns x.
template <int ROWS>
__device__ __forceinline__ void cheb_rows(const int32_t* __restrict__ nbi, const float* __restrict__ nbc,
                                          const int32_t* __restrict__ ovi, const float* __restrict__ ovc,
                                          const double* __restrict__ c, int K, int n, double* buf0,
                                          double* buf1, double (&out)[kChebRows]) {
    // neighbours 0 and 1 (local rows < 65536) packed in one register, neighbour 2 (or the
    // overflow marker) in another; ROWS rows per thread (n <= ROWS * kChebBlock)
    uint32_t i01[ROWS];
    int32_t i2[ROWS];
    float cf[ROWS][3];
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        const int r = threadIdx.x + j * kChebBlock;
        out[j] = 0.0;
        i01[j] = r < n ? ((uint32_t)nbi[3 * r] | ((uint32_t)nbi[3 * r + 1] << 16)) : 0u;
        i2[j] = r < n ? nbi[3 * r + 2] : 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) cf[j][d] = r < n ? nbc[3 * r + d] : 0.0f;
        if (r < n) out[j] = c[0] * buf0[r];
    }
    // the term before the current one is still in the buffer about to be written: the
    // recurrence t_{k+1} = 2 A~ t_k - t_{k-1} reads t_k and t_{k-1} of a row from LDS
    double* cur = buf0;
    double* nxt = buf1;
    for (int k = 1; k <= K; ++k) {
        const double ck = c[k];
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            const int r = threadIdx.x + j * kChebBlock;
            if (r < n) {
                const double tc = cur[r];
                double s = 0.0;
                if (cf[j][0] != 0.0f) s = fma((double)cf[j][0], cur[i01[j] & 0xFFFFu] - tc, s);
                if (cf[j][1] != 0.0f) s = fma((double)cf[j][1], cur[i01[j] >> 16] - tc, s);
                if (i2[j] >= 0) {
                    if (cf[j][2] != 0.0f) s = fma((double)cf[j][2], cur[i2[j]] - tc, s);
                } else {
                    for (int e = -i2[j] - 1; ovi[e] >= 0; ++e)
                        s = fma((double)ovc[e], cur[ovi[e]] - tc, s);
                }
                const double z = tc + s;
                const double tn = k == 1 ? z : 2.0 * z - nxt[r];
                nxt[r] = tn;
                out[j] = fma(ck, tn, out[j]);
            }
        }
        __syncthreads();
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    __syncthreads();
}

// Chebyshev products over cells laid out in preorder slots.  Thread t owns the J consecutive
// slots s = J t + j and keeps their last two terms in FP64 registers.  T_k = 2 A~ T_{k-1} - T_{k-2}
// is formed as
//   2 d T_{k-1}(s) + 2 a_l T_{k-1}(s-1) + 2 a_r T_{k-1}(s+1) + 2 a_o T_{k-1}(f_s) + 2 a_p T_{k-1}(f_{s-1})
//   - T_{k-2}(s),   d = 1 - (a_l + a_r + a_o + a_p) in FP64,
// with s - 1 and s + 1 taken from the thread's own registers or, across threads of a wave, from
// the neighbouring lane (DPP row shifts of the whole wave).  Every slot has one LDS read port:
// per term it reads the place f_s of the previous term vector.  A row takes its far neighbours
// from its own port and from the port of slot s - 1 (the thread's register, or lane l - 1's last
// port by DPP), so two far neighbours cost one read per slot; the layout's holes keep every row
// within these two ports.  A cell of one row group (J = 1) reads its two far neighbours itself
// instead (f, g: a second read per row, no DPP on the term's chain, fewer holes).  The FP32
// coefficients a = (2/rho) b are widened to FP64 and doubled exactly, so constant vectors stay
// constant to FP64 rounding.
//
// In the two term vectors, row groups 2p and 2p + 1 of thread t sit side by side (one 16-byte
// store per pair and term), an odd last row group alone (cheb_term_place).  x arrives at, and
// out[j] belongs to, the contract place j * 1024 + t of slot J t + j.
//
// A launch is built for at most MAXJ row groups per thread (MAXJ = 1 or kChebRows): its LDS holds
// two term vectors of BUF = MAXJ * 1024 doubles and the series coefficients (cells of more terms
// are read in natural order).  The one-group launch needs 24 KB and at most 64 VGPRs, so two
// workgroups share a CU.  Each cell's layout leaves its last slot (J * 1024 - 1) empty: unused
// ports read the zero kept at its term place in both vectors.
constexpr int kChebBuf = kChebBlock * kChebRows;
constexpr int kChebCoefCap = 1024;
constexpr int cheb_lds(int maxj) { return 2 * maxj * kChebBlock + kChebCoefCap; }
// the rational product: x, then q / z as complex, then up to 4095 level starts (doubles)
constexpr int rat_lds(int maxj) { return 3 * maxj * kChebBlock + 1024; }
// pre_far flag: the slot holds no row (a hole, or padding after the cell's last slot)
constexpr uint32_t kChebNoRow = 0x80000000u;

// Place of slot J t + j in a term vector.
__host__ __device__ constexpr int cheb_term_place(int t, int j, int J) {
    return (j >> 1) < (J >> 1) ? (j >> 1) * 2 * kChebBlock + 2 * t + (j & 1)
                               : (J >> 1) * 2 * kChebBlock + t;
}

// v of lane l - 1 (wave_shr:1; lane 0 receives 0) and of lane l + 1 (wave_shl:1; lane 63: 0).
__device__ __forceinline__ double lane_before(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double lane_after(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Registers of a thread's J slots; CT holds the doubled coefficients (FP64, or FP32 where the
// registers of 5 slots do not leave room).
template <int J, typename CT>
struct ChebRegs {
    uint32_t f[J];         // LDS byte offset (in a term vector) the slot's port reads
    uint32_t g;            // one row group: the second far neighbour's offset
    CT cl[J], cr[J], co[J], cp[J];
    double d2[J];          // 2 d
};

// Store the thread's J values into a term vector at their term places.
template <int J>
__device__ __forceinline__ void cheb_store_terms(double* v, const double (&u)[J]) {
#pragma unroll
    for (int p = 0; p < J / 2; ++p)
        *reinterpret_cast<double2*>(v + p * 2 * kChebBlock + 2 * threadIdx.x) = make_double2(u[2 * p], u[2 * p + 1]);
    if (J & 1) v[(J / 2) * 2 * kChebBlock + threadIdx.x] = u[J - 1];
}

// One term: t holds T_{k-1} of the thread's slots, u holds T_{k-2} and receives T_k (the two
// arrays swap roles from term to term, so no register is copied); the ports read buffer RD,
// T_k goes to the other buffer, out += ck T_k (ck = c_k on entry, c_knext on return), then a
// barrier.  FIRST: T_1 = A~ T_0.  ACTIVE (wave-uniform): the wave holds a slot in use; a wave
// without one only keeps the barriers.  LDS order: the next term's coefficient first, then the
// last row group's port (row 0 needs it, through lane l - 1), with odd J the one before it, then
// the others; rows are formed in their store units (with odd J the lone last row group first,
// then the pairs) and each unit is stored as soon as it is formed, since the last stores of the
// workgroup gate the barrier.
template <int J, typename CT, int RD, bool FIRST, int BUF>
__device__ __forceinline__ void cheb_term(double* lds, const ChebRegs<J, CT>& R, const double (&t)[J],
                                          double (&u)[J], bool active, double& ck, int knext,
                                          double (&out)[kChebRows]) {
    const double cnext = lds[2 * BUF + knext];  // the next term's coefficient, read ahead
    if (active) {
        const char* rd = reinterpret_cast<const char*>(lds + RD * BUF);
        double* wr = lds + (1 - RD) * BUF;
        double vf[J];
        vf[J - 1] = *reinterpret_cast<const double*>(rd + R.f[J - 1]);
        if (J & 1 && J > 1) vf[J - 2] = *reinterpret_cast<const double*>(rd + R.f[J - 2]);
#pragma unroll
        for (int j = 0; j < ((J & 1 && J > 1) ? J - 2 : J - 1); ++j) vf[j] = *reinterpret_cast<const double*>(rd + R.f[j]);
        const double before = lane_before(t[J - 1]);
        const double after = lane_after(t[0]);
        // one row group: the row's own second read; otherwise the previous slot's port
        const double port_before = J == 1 ? *reinterpret_cast<const double*>(rd + R.g)
                                          : lane_before(vf[J - 1]);
#pragma unroll
        for (int q = 0; q < J; ++q) {
            // odd J: the lone last row group first, then the pairs
            const int j = (J & 1) ? (q == 0 ? J - 1 : q - 1) : q;
            const double tl = j > 0 ? t[j - 1] : before;
            const double tr = j < J - 1 ? t[j + 1] : after;
            const double vp = j > 0 ? vf[j - 1] : port_before;
            double acc = FIRST ? R.d2[j] * t[j] : fma(R.d2[j], t[j], -u[j]);
            acc = fma((double)R.cl[j], tl, acc);
            acc = fma((double)R.cr[j], tr, acc);
            acc = fma((double)R.co[j], vf[j], acc);
            acc = fma((double)R.cp[j], vp, acc);
            u[j] = FIRST ? 0.5 * acc : acc;
            if ((j & 1) && (j >> 1) < (J >> 1)) {
                *reinterpret_cast<double2*>(wr + (j >> 1) * 2 * kChebBlock + 2 * threadIdx.x) = make_double2(u[j - 1], u[j]);
                asm volatile("" : "+v"(u[j - 1]), "+v"(u[j]) : : "memory");
            } else if (j == J - 1 && (J & 1)) {
                wr[(J / 2) * 2 * kChebBlock + threadIdx.x] = u[j];
                asm volatile("" : "+v"(u[j]) : : "memory");
            }
        }
#pragma unroll
        for (int j = 0; j < J; ++j) out[j] = fma(ck, u[j], out[j]);
    }
    ck = cnext;
    __syncthreads();
}

// The product for one preorder cell of `size` slots in J row groups (size <= 1024 J,
// K < kChebCoefCap).
// On entry lds[j * 1024 + t] = x of the row in slot J t + j (after a barrier); on return out[j]
// holds that row's result and the call has ended with a barrier.
template <int J, typename CT, int BUF, int NT>
__device__ __forceinline__ void cheb_pre_rows(const uint32_t* __restrict__ far, const float4* __restrict__ cf,
                                              const double* __restrict__ c, int K, int size, double* lds,
                                              double (&out)[kChebRows]) {
    constexpr int zero = cheb_term_place(kChebBlock - 1, J - 1, J);
    for (int k = threadIdx.x; k <= K; k += NT) lds[2 * BUF + k] = c[k];
    ChebRegs<J, CT> R;
    double x[J], y[J];
    const int wave_slot0 = __builtin_amdgcn_readfirstlane(threadIdx.x & ~(nw::kWave - 1)) * J;
    const bool active = wave_slot0 < size;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int q = threadIdx.x + j * kChebBlock;
        const uint32_t w = far[q];
        R.f[j] = 8u * (w & 0xFFFFu);
        if (J == 1) R.g = 8u * ((w >> 16) & 0x7FFFu);
        if (!(w & kChebNoRow)) {
            const float4 a = cf[q];
            R.cl[j] = (CT)(2.0f * a.x);
            R.cr[j] = (CT)(2.0f * a.y);
            R.co[j] = (CT)(2.0f * a.z);
            R.cp[j] = (CT)(2.0f * a.w);
            R.d2[j] = 2.0 * (1.0 - (((double)a.x + (double)a.y) + ((double)a.z + (double)a.w)));
            x[j] = lds[q];
        } else {  // hole or padding: zero, read by nobody but through the zero place
            R.cl[j] = R.cr[j] = R.co[j] = R.cp[j] = (CT)0;
            R.d2[j] = 0.0;
            x[j] = 0.0;
        }
        y[j] = 0.0;
    }
    __syncthreads();  // x read from the contract places: buffer 0 takes the term layout
    cheb_store_terms<J>(lds, x);
    if (threadIdx.x == 0) {  // (NT < 1024: the zero slot's thread does not exist)
        lds[zero] = 0.0;
        lds[BUF + zero] = 0.0;
    }
    __syncthreads();
    const double c0 = lds[2 * BUF];
#pragma unroll
    for (int j = 0; j < J; ++j) out[j] = c0 * x[j];
    // x = T_0, then y = T_1, x = T_2, y = T_3, ...
    // each term reads the next term's coefficient before its own ports (index clamped to K)
    double ck = lds[2 * BUF + 1];
    if (K >= 1) cheb_term<J, CT, 0, true, BUF>(lds, R, x, y, active, ck, min(2, K), out);
    int k = 2;
    for (; k + 1 <= K; k += 2) {
        cheb_term<J, CT, 1, false, BUF>(lds, R, y, x, active, ck, k + 1, out);
        cheb_term<J, CT, 0, false, BUF>(lds, R, x, y, active, ck, min(k + 2, K), out);
    }
    if (k <= K) cheb_term<J, CT, 1, false, BUF>(lds, R, y, x, active, ck, K, out);
}

// The product for one cell, preorder slots (base >= 0) or natural order, from x in LDS (at the
// row's LDS place) to out[] (the result of the row at LDS place threadIdx.x + j * 1024), in a
// launch built for MAXJ row groups (the store's cells need no more) and NT threads per
// workgroup.  NT = kChebSmallBlock serves stores whose cells are all preorder in one row group of
// at most NT slots (nw_propagator::cheb_small): the LDS places stay those of the 1024-thread
// layout, threads NT..1023 would own only empty slots and do not exist.  Returns the number of
// LDS places in use: out[j] is defined where threadIdx.x + j * 1024 is below it.
// The rational store's product for one cell (the contract of cheb_cell): x at the rows' level
// places in LDS (FP64, after a barrier); thread t owns the places t + j * 1024.  Per pole, the
// system (B - theta I) z = x by elimination over the tree: leaves to root
//   q_i = (x_i - sum_children B[i][c] q_c) / d_i,
// then root to leaves z_i = q_i - (B[i][parent] / d_i) z_parent, one level per barrier, q and z
// as FP64 complex in LDS (16 B per row); out = a0 x + 2 Re sum_j alpha_j z_j.  The level starts
// sit in LDS after the two term vectors.  Ends with a barrier; returns the places in use (n).
template <int MAXJ, int NT>
__device__ __forceinline__ int rat_cell(const ChebStore& cs, int64_t row0, int cell, int n, double* lds,
                                        double (&out)[kChebRows]) {
    // LDS: x (FP64, the caller's places) | q and z (FP64 complex) | level starts
    double2* q = reinterpret_cast<double2*>(lds + MAXJ * kChebBlock);
    uint16_t* lev = reinterpret_cast<uint16_t*>(lds + 3 * MAXJ * kChebBlock);
    const int D = cs.rat_depth[cell];
    const uint16_t* glev = cs.rat_lev + cs.rat_lev_off[cell];
    for (int l = threadIdx.x; l <= D; l += NT) lev[l] = glev[l];
    const uint32_t* link = cs.rat_link + row0;
    const float2* bb = cs.rat_b + row0;
    uint32_t lk[MAXJ];
    int lvl[MAXJ];
    float bdn[MAXJ], bc0[MAXJ], bc1[MAXJ];
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) {
        const int k = threadIdx.x + j * kChebBlock;
        lk[j] = 0u;
        bdn[j] = bc0[j] = bc1[j] = 0.0f;
        out[j] = 0.0;
        if (k < n) {
            out[j] = kRatDev[0] * lds[k];
            lk[j] = link[k];
            bdn[j] = bb[k].y;
            const int fc = (int)((lk[j] >> 13) & 0x1FFFu), cnt = (int)(lk[j] >> 26);
            if (cnt > 0) bc0[j] = bb[fc].x;
            if (cnt > 1) bc1[j] = bb[fc + 1].x;
        }
    }
    __syncthreads();
    // the level of every owned row (-1: no row), by bisection of the level starts
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) {
        const int k = threadIdx.x + j * kChebBlock;
        int a = 0, b = D;  // lev[a] <= k < lev[b]
        while (b - a > 1) {
            const int m = (a + b) >> 1;
            if ((int)lev[m] <= k) a = m;
            else b = m;
        }
        lvl[j] = k < n ? a : -1;
    }
    const double2* gid = cs.rat_invd + row0;
    for (int pj = 0; pj < kRatPoles; ++pj) {
        const double al_re2 = 2.0 * kRatDev[3 + 4 * pj], al_im2 = 2.0 * kRatDev[4 + 4 * pj];
        double2 id[MAXJ];
#pragma unroll
        for (int j = 0; j < MAXJ; ++j) {
            const int k = threadIdx.x + j * kChebBlock;
            id[j] = k < n ? gid[(int64_t)pj * cs.num_rows + k] : make_double2(0.0, 0.0);
        }
        // leaves to root
        for (int l = D - 1; l >= 0; --l) {
#pragma unroll
            for (int j = 0; j < MAXJ; ++j) {
                const int k = threadIdx.x + j * kChebBlock;
                if (lvl[j] == l) {
                    double re = lds[k], im = 0.0;
                    const int fc = (int)((lk[j] >> 13) & 0x1FFFu), cnt = (int)(lk[j] >> 26);
                    if (cnt > 0) {
                        const double2 c0 = q[fc];
                        re = fma(-(double)bc0[j], c0.x, re);
                        im = fma(-(double)bc0[j], c0.y, im);
                    }
                    if (cnt > 1) {
                        const double2 c1 = q[fc + 1];
                        re = fma(-(double)bc1[j], c1.x, re);
                        im = fma(-(double)bc1[j], c1.y, im);
                    }
                    for (int c = 2; c < cnt; ++c) {  // more than two children (a soma)
                        const double b = bb[fc + c].x;
                        const double2 cc = q[fc + c];
                        re = fma(-b, cc.x, re);
                        im = fma(-b, cc.y, im);
                    }
                    q[k] = make_double2(fma(re, id[j].x, -im * id[j].y), fma(re, id[j].y, im * id[j].x));
                }
            }
            __syncthreads();
        }
        // root to leaves
        for (int l = 0; l < D; ++l) {
#pragma unroll
            for (int j = 0; j < MAXJ; ++j) {
                const int k = threadIdx.x + j * kChebBlock;
                if (lvl[j] == l) {
                    double2 z = q[k];
                    if (l > 0) {
                        const double2 zp = q[lk[j] & 0x1FFFu];
                        const double er = (double)bdn[j] * id[j].x, ei = (double)bdn[j] * id[j].y;
                        z.x -= fma(er, zp.x, -ei * zp.y);
                        z.y -= fma(er, zp.y, ei * zp.x);
                    }
                    q[k] = z;
                    out[j] = fma(al_re2, z.x, out[j]);
                    out[j] = fma(-al_im2, z.y, out[j]);
                }
            }
            __syncthreads();
        }
    }
    return n < MAXJ * kChebBlock ? n : MAXJ * kChebBlock;
}

// KIND selects the products a kernel is built with: kProdAny (the store's flag decides at run
// time), kProdRational or kProdChebyshev (launches for one store only: no registers or code of
// the other product).
constexpr int kProdAny = 0, kProdRational = 1, kProdChebyshev = 2;
// LDS (doubles) of a launch built for KIND and MAXJ
constexpr int prod_lds(int maxj, int kind) {
    return kind == kProdChebyshev ? cheb_lds(maxj)
         : kind == kProdRational ? rat_lds(maxj)
         : (cheb_lds(maxj) > rat_lds(maxj) ? cheb_lds(maxj) : rat_lds(maxj));
}
template <int MAXJ, int NT = kChebBlock, int KIND = kProdAny>
__device__ __forceinline__ int cheb_cell(const ChebStore& cs, int64_t row0, int cell, int n, int64_t base,
                                         double* lds, double (&out)[kChebRows]) {
    static_assert(NT == kChebBlock || MAXJ == 1, "small workgroups serve one row group");
    if constexpr (KIND == kProdRational) {
        return rat_cell<MAXJ, NT>(cs, row0, cell, n, lds, out);
    } else {
    if (KIND == kProdAny && cs.rational) return rat_cell<MAXJ, NT>(cs, row0, cell, n, lds, out);
    constexpr int BUF = MAXJ * kChebBlock;
    const double* c = cs.c + cs.off[cell];
    const int K = cs.K[cell];
    if constexpr (NT == kChebBlock) {
        if (base < 0) {
            cheb_rows<MAXJ>(cs.nb_idx + 3 * row0, cs.nb_coef + 3 * row0, cs.ov_idx, cs.ov_coef, c,
                            K, n, lds, lds + BUF, out);
            return n;
        }
    }
    const uint32_t* far = cs.pre_far + base;
    const float4* cf = cs.pre_coef + base;
    const int size = cs.pre_size[cell] & 0xFFFF;
    const int J = cs.pre_size[cell] >> 16;
    if constexpr (MAXJ == 1) {
        cheb_pre_rows<1, double, BUF, NT>(far, cf, c, K, size, lds, out);
        return kChebBlock;  // J = 1 in a one-group launch: out[1..] are never read
    } else {
        if (J <= 1) cheb_pre_rows<1, double, BUF, NT>(far, cf, c, K, size, lds, out);
        else if (J == 2) cheb_pre_rows<2, double, BUF, NT>(far, cf, c, K, size, lds, out);
        else if (J == 3) cheb_pre_rows<3, double, BUF, NT>(far, cf, c, K, size, lds, out);
        else if (J == 4) cheb_pre_rows<4, double, BUF, NT>(far, cf, c, K, size, lds, out);
        else cheb_pre_rows<5, double, BUF, NT>(far, cf, c, K, size, lds, out);
    }
    return J * kChebBlock;
    }
}
