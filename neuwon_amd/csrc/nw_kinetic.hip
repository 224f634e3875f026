// LTI kinetic advance grouped by voltage bucket (replaces the kernel generated by
// lti_sim/codegen.py:39-106 for schemes with one linear input, the membrane voltage).
//
// The reference kernel interpolates each instance's propagator between the two table nodes of its
// bucket and multiplies it into the instance's state.  Every instance of a wavefront that shares
// a bucket needs the same two node matrices, so the kernel reads them once per wavefront and
// bucket:
//   x' = y0 + f (y1 - y0),  y0 = T_b x,  y1 = T_(b+1) x      (codegen.py:81-94, linear terms)
// as FP64 FMAs with the node entries in SGPRs (scalar loads: no per-lane table traffic), each
// entry shared by the several instances a lane holds.  A wavefront whose instances fall into more
// than `group_max` distinct buckets gathers each instance's node matrices itself (16-byte loads
// from the padded FP32 copy).  Both forms do the same FP64 operations in the same order, so their
// results are bitwise equal.
#include <algorithm>
#include <cstring>

#include "nw_internal.h"

struct nw_lti_table {
    int device = 0;
    int32_t S = 0;
    int32_t num_nodes = 0;
    float vmin = 0.0f, vmax = 0.0f;
    int32_t pad_stride = 0;        // floats per node of f32_pad (S*S rounded up to 4)
    const float* f32 = nullptr;    // caller's table, [num_nodes][S][S]
    float* f32_pad = nullptr;      // [num_nodes][pad_stride], zero padded
    double* f64 = nullptr;         // [num_nodes][S][S], the FP32 values widened
    int num_cus = 0;               // compute units of the device
};

namespace nw {

// Widen the caller's FP32 node table to FP64 and copy it to the padded FP32 layout.
__global__ void k_lti_table_layouts(int64_t num_nodes, int S, int pad, const float* __restrict__ src,
                                    float* __restrict__ f32_pad, double* __restrict__ f64) {
    const int64_t total = num_nodes * (int64_t)pad;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t node = e / pad;
        const int q = (int)(e - node * pad);
        const float x = q < S * S ? src[node * S * S + q] : 0.0f;
        f32_pad[e] = x;
        if (q < S * S) f64[node * S * S + q] = (double)x;
    }
}

using ConstF64 = const __attribute__((address_space(4))) double*;
typedef double f64x4 __attribute__((ext_vector_type(4)));

struct LtiGroupArgs {
    int64_t M;
    int32_t num_nodes;
    int32_t group_max;
    double vmin, bucket_frq;
    const float* f32_pad;   // padded FP32 nodes (gather path)
    int32_t pad;
    const double* f64;      // FP64 nodes (grouped path)
    const float* voltage;
    const int32_t* seg_index;
    float* states;
    int64_t stride;
    float conserve_sum;
    const float* weights;
    const float* gmax;
    float* cond_out;
};

// Wavefronts whose instances fall into at most this many distinct buckets take the grouped path.
constexpr int kLtiGroupMax = 2;

// Instances per lane in k_lti_grouped: every node entry a grouped wavefront reads by a scalar load
// feeds this many FMAs per lane (the load's latency spread over more work), within the register
// budget of the states held per instance (two FP64 vectors of S).
template <int S>
constexpr int lti_instances_per_lane() { return 1; }

// out[p] = y0 + f[p] (y1 - y0), y0 = T0 x[p], y1 = T1 x[p] for the P instances of the lane, T0
// and T1 the node matrices at `t0` (T1 at t0 + S*S): entries in row-major order, every row's two
// sums over the columns in ascending order (the sums of the reference kernel, codegen.py:81-94),
// in FP64, each entry used for all P instances as it arrives.
template <int S, int P, typename Ptr>
__device__ __forceinline__ void interp_products(Ptr t0, const double (&x)[P][S], const double (&f)[P],
                                                double (&out)[P][S]) {
#pragma unroll
    for (int r = 0; r < S; ++r) {
        double y0[P], y1[P];
#pragma unroll
        for (int p = 0; p < P; ++p) y0[p] = y1[p] = 0.0;
#pragma unroll
        for (int c = 0; c < S; ++c) {
            const double a0 = (double)t0[r * S + c], a1 = (double)t0[S * S + r * S + c];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                y0[p] = fma(a0, x[p][c], y0[p]);
                y1[p] = fma(a1, x[p][c], y1[p]);
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) out[p][r] = fma(f[p], y1[p] - y0[p], y0[p]);
    }
}

// The gather form: the lane's own node matrices for one instance, 16-byte loads of the padded
// FP32 nodes consumed in the row-major order of interp_products (the same operations in the same
// order, so both forms give bitwise equal results).
template <int S>
__device__ __forceinline__ void gather_product(const LtiGroupArgs& g, int b, double f,
                                               const double (&x)[S], double (&out)[S]) {
    const float4* t4 = reinterpret_cast<const float4*>(g.f32_pad + (int64_t)b * g.pad);
    const int q1 = g.pad / 4;
    double y0 = 0.0, y1 = 0.0;
#pragma unroll
    for (int q = 0; q < (S * S + 3) / 4; ++q) {
        const float4 a0 = t4[q], a1 = t4[q1 + q];
        const float e0[4] = {a0.x, a0.y, a0.z, a0.w}, e1[4] = {a1.x, a1.y, a1.z, a1.w};
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int e = 4 * q + jj;
            if (e < S * S) {
                y0 = fma((double)e0[jj], x[e % S], y0);
                y1 = fma((double)e1[jj], x[e % S], y1);
                if (e % S == S - 1) {
                    out[e / S] = fma(f, y1 - y0, y0);
                    y0 = y1 = 0.0;
                }
            }
        }
    }
}

// Every 64-byte line of the two node matrices at t0 requested at once before a pass reads them
// (one scalar-cache miss latency for the pass instead of one per dependent scalar load): lines
// from t0 rounded down to 64 bytes, in groups of 8 (the FP64 table carries 8 KB of padding for
// the last groups).  The loads complete inside the asm statement.
#define NW_TL(off) "s_load_dword %0, %1, " off "\n"
#define NW_T8(a, b) NW_TL(a "00") NW_TL(a "40") NW_TL(a "80") NW_TL(a "c0") \
                    NW_TL(b "00") NW_TL(b "40") NW_TL(b "80") NW_TL(b "c0")
#define NW_TG1 NW_T8("0x0", "0x1")
#define NW_TG2 NW_TG1 NW_T8("0x2", "0x3")
#define NW_TG3 NW_TG2 NW_T8("0x4", "0x5")
#define NW_TG4 NW_TG3 NW_T8("0x6", "0x7")
#define NW_TG5 NW_TG4 NW_T8("0x8", "0x9")
#define NW_TG6 NW_TG5 NW_T8("0xa", "0xb")
#define NW_TG7 NW_TG6 NW_T8("0xc", "0xd")
#define NW_TG8 NW_TG7 NW_T8("0xe", "0xf")
#define NW_TG9 NW_TG8 NW_T8("0x10", "0x11")
#define NW_TOUCH(lines) asm volatile(lines "s_waitcnt lgkmcnt(0)" : "=&s"(d) : "s"(a) : "memory")

template <int S>
__device__ __forceinline__ void touch_node_pair(ConstF64 t0) {
    constexpr int kLines = (2 * S * S * 8 + 63) / 64 + 1;
    constexpr int kGroups = (kLines + 7) / 8;
    static_assert(kGroups <= 9, "node pair beyond the touch groups");
    const ConstF64 a = (ConstF64)((uintptr_t)t0 & ~(uintptr_t)63);
    int d;
    if constexpr (kGroups == 1) NW_TOUCH(NW_TG1);
    else if constexpr (kGroups == 2) NW_TOUCH(NW_TG2);
    else if constexpr (kGroups == 3) NW_TOUCH(NW_TG3);
    else if constexpr (kGroups == 4) NW_TOUCH(NW_TG4);
    else if constexpr (kGroups == 5) NW_TOUCH(NW_TG5);
    else if constexpr (kGroups == 6) NW_TOUCH(NW_TG6);
    else if constexpr (kGroups == 7) NW_TOUCH(NW_TG7);
    else if constexpr (kGroups == 8) NW_TOUCH(NW_TG8);
    else NW_TOUCH(NW_TG9);
    asm volatile("" ::"s"(d));
}
#undef NW_TOUCH
#undef NW_TG9
#undef NW_TG8
#undef NW_TG7
#undef NW_TG6
#undef NW_TG5
#undef NW_TG4
#undef NW_TG3
#undef NW_TG2
#undef NW_TG1
#undef NW_T8
#undef NW_TL
// Schemes from this many states touch their node pair before a pass (smaller pairs span few
// lines, which the pass's own loads fetch about as fast).
constexpr int kTouchMinStates = 6;

// A wavefront-uniform value made opaque in SGPRs (the row stride: recomputed row bases then stay
// out of the SGPRs the node entries need), and a wavefront-uniform pointer likewise: an access
// p[o] with a 32-bit lane offset then takes the SGPR-base form (one VGPR per address instead of a
// 64-bit VGPR pair per array row).
template <typename T>
__device__ __forceinline__ T opaque(T v) {
    asm volatile("" : "+s"(v));
    return v;
}

template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* sbase(T* p) {
    auto q = (__attribute__((address_space(1))) T*)p;
    asm volatile("" : "+s"(q));
    return q;
}

// CONSERVE (codegen.py:95-102), the states and the conductance of instance base + o from out
// (every address a wavefront-uniform base plus the lane's 32-bit offset); gmax: the instance's
// maximal conductance (read with the tile's states, before its stores).  The weights are read
// through the constant address space (scalar loads: the kernel never writes them).
template <int S, bool COND>
__device__ __forceinline__ void finish_lane(const LtiGroupArgs& g, int64_t base, uint32_t o,
                                            double (&out)[S], float gmax) {
    if (g.conserve_sum > 0.0f) {
        double tot = 0.0;
#pragma unroll
        for (int r = 0; r < S; ++r) tot += out[r];
        const double corr = (double)g.conserve_sum / tot;
#pragma unroll
        for (int r = 0; r < S; ++r) out[r] *= corr;
    }
    const int64_t stride = opaque(g.stride);
#pragma unroll
    for (int k = 0; k < S; ++k) sbase(g.states + k * stride + base)[o] = (float)out[k];
    if constexpr (COND) {
        double w = 0.0;
        const auto wt = (const __attribute__((address_space(4))) float*)g.weights;
#pragma unroll
        for (int k = 0; k < S; ++k) w = fma((double)wt[k], (double)(float)out[k], w);
        sbase(g.cond_out + base)[o] = (float)((double)gmax * w);
    }
}

// The bucket of the first lane left in rem (instance sets in order), or -1 when all are empty.
// rem holds ballots, so the scan is uniform over the wavefront.
template <int P>
__device__ __forceinline__ int next_bucket(const uint64_t (&rem)[P], const int (&b)[P]) {
#pragma unroll
    for (int p = 0; p < P; ++p)
        if (rem[p]) return __builtin_amdgcn_readlane(b[p], __builtin_ctzll(rem[p]));
    return -1;
}

// Distinct buckets among the wavefront's active instances, counted up to cap + 1.
template <int P>
__device__ __forceinline__ int distinct_buckets(const uint64_t (&lanes)[P], const int (&b)[P],
                                                int cap) {
    uint64_t scan[P];
#pragma unroll
    for (int p = 0; p < P; ++p) scan[p] = lanes[p];
    int distinct = 0;
    for (; distinct <= cap; ++distinct) {
        const int u = next_bucket<P>(scan, b);
        if (u < 0) break;
#pragma unroll
        for (int p = 0; p < P; ++p) scan[p] &= ~__ballot(b[p] == u);
    }
    return distinct;
}

// Tiles of `tile` consecutive instances; the G workgroups take balanced runs of consecutive tiles,
// so a workgroup's wavefronts meet the same few buckets from one tile to the next (instances in
// segment order: neighbouring tiles sit on neighbouring voltages).  Workgroups are dispatched to
// the device's 8 XCDs round robin (workgroup i on XCD i mod 8); run (i mod 8) G/8 + i/8 gives
// each XCD one contiguous stretch of instances, whose node matrices then stay in that XCD's L2.
__device__ __forceinline__ void tile_range(const LtiGroupArgs& g, int tile, int64_t& begin,
                                           int64_t& end) {
    const int64_t tiles = (g.M + tile - 1) / tile, G = gridDim.x, i = blockIdx.x;
    const int64_t r = G % 8 == 0 ? (i % 8) * (G / 8) + i / 8 : i;
    begin = r * tiles / G;
    end = (r + 1) * tiles / G;
}

// One wavefront per 64 P consecutive instances of a tile (lane l holds instances l, l + 64, ...,
// so every load and store is 64 consecutive words); loads run one tile ahead, every one
// unconditional (past the end of the run or of the instances they repeat the last valid
// address), so they stay in flight across the current tile's arithmetic.  A wavefront of at most
// group_max distinct buckets makes one pass per bucket: the bucket's two node matrices by scalar
// loads (FP64 FMAs with SGPR operands, shared by the lane's P instances), the instances of the
// bucket taking the result; other wavefronts gather per instance.
template <int S, int P, bool COND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_lti_grouped(LtiGroupArgs g) {
    constexpr int kTile = 256 * P;
    const int nb = g.num_nodes - 1;
    int64_t tile, tile_end;
    tile_range(g, kTile, tile, tile_end);
    if (tile >= tile_end) return;
    const uint32_t lane0 = (threadIdx.x >> 6) * 64 * P + (threadIdx.x & 63);
    float v_n[P], x_n[P][S], gm_n[P];
    auto load = [&](int64_t t) {
        const int64_t stride = opaque(g.stride);
        const int64_t base = (t < tile_end ? t : tile_end - 1) * kTile;
        const uint32_t last = (uint32_t)std::min<int64_t>(g.M - 1 - base, kTile - 1);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const uint32_t o = std::min<uint32_t>(lane0 + 64 * p, last);
            v_n[p] = g.seg_index ? g.voltage[(uint32_t)sbase(g.seg_index + base)[o]]
                                : sbase(g.voltage + base)[o];
#pragma unroll
            for (int k = 0; k < S; ++k) x_n[p][k] = sbase(g.states + k * stride + base)[o];
            gm_n[p] = COND ? sbase(g.gmax + base)[o] : 0.0f;
        }
    };
    load(tile);
    for (; tile < tile_end; ++tile) {
        const int64_t base = tile * kTile;
        const int64_t left = g.M - base;
        bool active[P];
        int b[P];
        double f[P], x[P][S];
        float gm[P];
        uint64_t lanes[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            gm[p] = gm_n[p];
            active[p] = (int64_t)(lane0 + 64 * p) < left;
            b[p] = -1;
            f[p] = 0.0;
#pragma unroll
            for (int k = 0; k < S; ++k) x[p][k] = (double)x_n[p][k];
            if (active[p]) bucket_of<double>((double)v_n[p], g.vmin, g.bucket_frq, nb, b[p], f[p]);
            lanes[p] = __ballot(active[p]);
        }
        load(tile + 1);
        if (distinct_buckets<P>(lanes, b, g.group_max) <= g.group_max) {
            uint64_t rem[P];
            double fin[P][S];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                rem[p] = lanes[p];
#pragma unroll
                for (int k = 0; k < S; ++k) fin[p][k] = 0.0;
            }
            for (;;) {
                const int u = next_bucket<P>(rem, b);
                if (u < 0) break;
                bool mine[P], any = false;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    mine[p] = b[p] == u;
                    rem[p] &= ~__ballot(mine[p]);
                    any = any || mine[p];
                }
                // the address from a copy of u the compiler cannot equate with the lanes' own b
                // (where they are equal, an address in b would be per lane)
                int ua = u;
                asm volatile("" : "+s"(ua));
                if (any) {
                    // every instance lies in one bucket: it is finished in that bucket's pass
                    double out[P][S];
                    const ConstF64 t0 = (ConstF64)(g.f64 + (int64_t)ua * S * S);
                    if constexpr (S >= kTouchMinStates) touch_node_pair<S>(t0);
                    interp_products<S, P>(t0, x, f, out);
                    // the products complete here: sunk into the per-instance branches below, they
                    // would hold every node entry in SGPRs across the first branch
#pragma unroll
                    for (int p = 0; p < P; ++p)
#pragma unroll
                        for (int k = 0; k < S; ++k) asm volatile("" : "+v"(out[p][k]));
#pragma unroll
                    for (int p = 0; p < P; ++p)
#pragma unroll
                        for (int k = 0; k < S; ++k) fin[p][k] = mine[p] ? out[p][k] : fin[p][k];
                }
            }
            // every lane's result stored once, after the passes, with no branch around the
            // stores (lanes past the instances are masked)
#pragma unroll
            for (int p = 0; p < P; ++p)
                if (active[p]) finish_lane<S, COND>(g, base, lane0 + 64 * p, fin[p], gm[p]);
        } else {
#pragma unroll
            for (int p = 0; p < P; ++p)
                if (active[p]) {
                    double out[S];
                    gather_product<S>(g, b[p], f[p], x[p], out);
                    finish_lane<S, COND>(g, base, lane0 + 64 * p, out, gm[p]);
                }
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Batched FP64 matrix exponential of small rate matrices (the table build of
// lti_sim/lti_model.py:51-65: P = expm(A dt), every column divided by its sum).
// One wavefront per matrix, padded to 16 x 16 in LDS; products on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64, four per product).  A rate matrix has non-negative off-diagonal
// entries, so X = A dt + lambda I (lambda = max |A_ii dt|) is non-negative and
// expm(A dt) = e^-lambda expm(X): with X scaled by 2^-s to a 1-norm of at most 1/2, the Taylor
// terms of expm(X 2^-s) never cancel (18 terms by Horner's rule, remainder below 1e-22), and s
// squarings of e^(-lambda 2^-s) expm(X 2^-s) follow, renormalised column by column when the
// scheme conserves its total and normalised columns are asked for.
// ----------------------------------------------------------------------------------------------
// C = A B for 16 x 16 row-major FP64 matrices in LDS, one wavefront; C may alias A or B.
// Operands of v_mfma_f64_16x16x4f64 per lane l: A[l % 16][k0 + l / 16], B[k0 + l / 16][l % 16];
// result register r of lane l: row 4 r + l / 16, column l % 16 (measured on the device with unit
// operands, lane by lane).
__device__ __forceinline__ void mm16(const double* A, const double* B, double* C) {
    const int l = threadIdx.x & 63, j = l & 15, q = l >> 4;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const int k = 4 * kk + q;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[j * 16 + k], B[k * 16 + j], acc, 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) C[(4 * r + q) * 16 + j] = acc[r];
    __syncthreads();
}

constexpr int kExpmTerms = 18;

__global__ void __launch_bounds__(64) k_expm_batched(int64_t count, int S, const double* __restrict__ A,
                                                     double dt, int normalise, double* __restrict__ P) {
    __shared__ double X[256], R[256];
    __shared__ double red[64];
    const int64_t m = blockIdx.x;
    if (m >= count) return;
    const int l = threadIdx.x;
    const double* a = A + m * S * S;
    for (int e = l; e < 256; e += 64) {
        const int r = e >> 4, c = e & 15;
        X[e] = (r < S && c < S) ? a[r * S + c] * dt : 0.0;
    }
    __syncthreads();
    // lambda = max |X_ii|; the 1-norm of X + lambda I (column sums of absolute values)
    red[l] = l < S ? fabs(X[l * 17]) : 0.0;
    __syncthreads();
    double lam = 0.0;
    for (int k = 0; k < 16; ++k) lam = fmax(lam, red[k]);
    __syncthreads();
    double col = 0.0;
    if (l < 16)
        for (int r = 0; r < 16; ++r) col += fabs(X[r * 16 + l] + (r == l && l < S ? lam : 0.0));
    red[l] = col;
    __syncthreads();
    double norm = 0.0;
    for (int k = 0; k < 16; ++k) norm = fmax(norm, red[k]);
    // conservative: every column of A dt sums to zero up to its rounding
    if (l < S) {
        double sum = 0.0, mag = 0.0;
        for (int r = 0; r < S; ++r) {
            sum += X[r * 16 + l];
            mag += fabs(X[r * 16 + l]);
        }
        red[l] = fabs(sum) <= 64.0 * 2.220446049250313e-16 * mag ? 1.0 : 0.0;
    }
    __syncthreads();
    bool conservative = true;
    for (int k = 0; k < S; ++k) conservative = conservative && red[k] != 0.0;
    __syncthreads();
    int s = 0;
    while (s < 1000 && ldexp(norm, -s) > 0.5) ++s;
    const double scale = ldexp(1.0, -s);
    __syncthreads();
    for (int e = l; e < 256; e += 64) {
        const int r = e >> 4, c = e & 15;
        X[e] = (X[e] + (r == c && r < S ? lam : 0.0)) * scale;
        R[e] = r == c ? 1.0 : 0.0;
    }
    __syncthreads();
    // R = I + X/1 (I + X/2 (... (I + X/18)))
    for (int k = kExpmTerms; k >= 1; --k) {
        mm16(X, R, R);
        const double inv = 1.0 / (double)k;
        for (int e = l; e < 256; e += 64) R[e] = R[e] * inv + ((e >> 4) == (e & 15) ? 1.0 : 0.0);
        __syncthreads();
    }
    const double shrink = exp(-lam * scale);
    for (int e = l; e < 256; e += 64) R[e] *= shrink;
    __syncthreads();
    // every column sum of the scaled R on red[0..S-1]
    auto column_sums = [&]() {
        if (l < S) {
            double t = 0.0;
            for (int r = 0; r < S; ++r) t += R[r * 16 + l];
            red[l] = t;
        }
        __syncthreads();
    };
    // A scheme whose columns sum to zero (it conserves its total) has column-stochastic
    // propagators at every squaring stage, expm(A dt 2^(k - s)): with normalised output their
    // columns are rescaled to sum one after each squaring, which removes the error the squarings
    // would otherwise grow along that direction.
    for (int k = 0; k < s; ++k) {
        mm16(R, R, R);
        if (normalise && conservative) {
            column_sums();
            for (int e = l; e < 256; e += 64)
                if ((e & 15) < S) R[e] /= red[e & 15];
            __syncthreads();
        }
    }
    if (normalise) column_sums();
    double* p = P + m * S * S;
    for (int e = l; e < S * S; e += 64) {
        const int r = e / S, c = e % S;
        p[e] = normalise ? R[r * 16 + c] / red[c] : R[r * 16 + c];
    }
}

}  // namespace nw

using namespace nw;

extern "C" {

int nw_lti_expm_batched(int64_t count, int32_t S, const double* A, double time_step,
                        int32_t normalise_columns, double* P, void* stream) {
    NW_REQUIRE(S >= 1 && S <= 16, "state count %d not supported (1..16)", S);
    NW_REQUIRE(count >= 0 && count <= 0x7FFFFFFF, "bad matrix count");
    if (count == 0) return NW_OK;
    NW_REQUIRE(A && P, "nw_lti_expm_batched: null argument");
    NW_REQUIRE(time_step >= 0.0, "negative time step");
    k_expm_batched<<<(unsigned)count, 64, 0, (hipStream_t)stream>>>(count, S, A, time_step,
                                                                     normalise_columns ? 1 : 0, P);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_lti_table_create(int device, int32_t S, int32_t num_nodes, float vmin, float vmax,
                        const float* node_table, nw_lti_table** out) {
    NW_REQUIRE(out && node_table, "nw_lti_table_create: null argument");
    NW_REQUIRE(S >= 1 && S <= 16, "state count %d not supported (1..16)", S);
    NW_REQUIRE(num_nodes >= 2 && vmax > vmin, "bad table grid");
    *out = nullptr;
    NW_HIP(hipSetDevice(device));
    int cus = 0;
    NW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    nw_lti_table* t = new nw_lti_table;
    t->num_cus = cus;
    t->device = device;
    t->S = S;
    t->num_nodes = num_nodes;
    t->vmin = vmin;
    t->vmax = vmax;
    t->pad_stride = (S * S + 3) / 4 * 4;
    t->f32 = node_table;
    // the FP64 nodes are followed by 8 KB of zeros, the lines touch_node_pair requests past the
    // last node pair
    const size_t n_pad = (size_t)num_nodes * t->pad_stride, n64 = (size_t)num_nodes * S * S + 1024;
    hipError_t e = hipMalloc(&t->f32_pad, sizeof(float) * n_pad);
    if (e == hipSuccess) e = hipMalloc(&t->f64, sizeof(double) * n64);
    if (e == hipSuccess) e = hipMemset(t->f64, 0, sizeof(double) * n64);
    if (e == hipSuccess) {
        k_lti_table_layouts<<<256, 256>>>(num_nodes, S, t->pad_stride, node_table, t->f32_pad,
                                          t->f64);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (t->f32_pad) (void)hipFree(t->f32_pad);
        if (t->f64) (void)hipFree(t->f64);
        delete t;
        NW_HIP(e);
    }
    *out = t;
    return NW_OK;
}

int nw_lti_table_destroy(nw_lti_table* t) {
    if (!t) return NW_OK;
    if (t->f32_pad) (void)hipFree(t->f32_pad);
    if (t->f64) (void)hipFree(t->f64);
    delete t;
    return NW_OK;
}

int nw_lti_table_advance(const nw_lti_table* t, int64_t M, const float* voltage,
                         const int32_t* seg_index, float* states, int64_t state_stride,
                         float conserve_sum, int32_t n_steps, const float* weights,
                         const float* gmax, float* cond_out, int32_t group_max, void* stream) {
    NW_REQUIRE(t, "nw_lti_table_advance: null table");
    NW_REQUIRE(M >= 0 && state_stride >= M && n_steps >= 0, "bad sizes");
    NW_REQUIRE(group_max >= -1 && group_max <= 64, "group_max must be -1 (default) or 0..64");
    if (M == 0 || n_steps == 0) return NW_OK;
    NW_REQUIRE(voltage && states, "nw_lti_table_advance: null argument");
    NW_REQUIRE(!cond_out || (weights && gmax), "conductance output needs weights and gmax");
    if (n_steps > 1)   // the input held over the launch: the interpolated propagator in registers
        return nw_lti_advance(t->S, t->num_nodes, t->vmin, t->vmax, t->f32, M, voltage, seg_index,
                              states, state_stride, conserve_sum, n_steps, weights, gmax, cond_out,
                              stream);
    LtiGroupArgs g{};
    g.M = M;
    g.num_nodes = t->num_nodes;
    g.group_max = group_max < 0 ? kLtiGroupMax : group_max;
    g.vmin = (double)t->vmin;
    g.bucket_frq = (double)(t->num_nodes - 1) / ((double)t->vmax - (double)t->vmin);
    g.f32_pad = t->f32_pad;
    g.pad = t->pad_stride;
    g.f64 = t->f64;
    g.voltage = voltage;
    g.seg_index = seg_index;
    g.states = states;
    g.stride = state_stride;
    g.conserve_sum = conserve_sum;
    g.weights = weights;
    g.gmax = gmax;
    g.cond_out = cond_out;
    // as many workgroups as are resident at once on the device, each a run of consecutive tiles
    // (one round: no tail of late workgroups)
    hipStream_t s = (hipStream_t)stream;
    auto launch = [&](auto kernel, int tile) -> int {
        int per_cu = 0;
        NW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
        const int64_t resident = std::max<int64_t>(1, (int64_t)std::max(per_cu, 1) * t->num_cus);
        const unsigned grid = (unsigned)std::min<int64_t>((M + tile - 1) / tile, resident);
        kernel<<<grid, 256, 0, s>>>(g);
        return NW_OK;
    };
    int rc = NW_OK;
#define NW_LTI_G(K)                                                               \
    case K: {                                                                     \
        constexpr int P = lti_instances_per_lane<K>();                            \
        rc = cond_out ? launch(k_lti_grouped<K, P, true>, 256 * P)                \
                      : launch(k_lti_grouped<K, P, false>, 256 * P);              \
        break;                                                                    \
    }
    switch (t->S) {
        NW_LTI_G(1) NW_LTI_G(2) NW_LTI_G(3) NW_LTI_G(4) NW_LTI_G(5) NW_LTI_G(6) NW_LTI_G(7)
        NW_LTI_G(8) NW_LTI_G(9) NW_LTI_G(10) NW_LTI_G(11) NW_LTI_G(12) NW_LTI_G(13)
        NW_LTI_G(14) NW_LTI_G(15) NW_LTI_G(16)
        default: NW_REQUIRE(false, "state count %d not supported (1..16)", t->S);
    }
#undef NW_LTI_G
    if (rc) return rc;
    NW_HIP(hipGetLastError());
    return NW_OK;
}

}  // extern "C"
