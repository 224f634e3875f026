// LTI kinetic advance grouped by voltage bucket (replaces the kernel generated by
// lti_sim/codegen.py:39-106 for schemes with one linear input, the membrane voltage).
//
// The reference kernel interpolates each instance's propagator between the two table nodes of its
// bucket and multiplies it into the instance's state.  Every instance of a wave that shares a
// bucket needs the same two node matrices, so the kernel reads them once per wave and bucket:
//   x' = y0 + f (y1 - y0),  y0 = T_b x,  y1 = T_(b+1) x      (codegen.py:81-94, linear terms)
// with T_b read by scalar loads (wave-uniform addresses: SGPR operands of the FP64 FMAs, no
// per-lane table traffic).  A wave whose lanes fall into more than `group_max` distinct buckets
// gathers each lane's node matrices itself (16-byte loads from the padded FP32 copy).  Both paths
// do the same FP64 operations in the same order on the same values, so an instance's result does
// not depend on the path its wave took.
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
    int64_t tiles_per_group;
};

// Waves whose lanes fall into at most this many distinct buckets take the grouped path.
constexpr int kLtiGroupMax = 2;

// out = y0 + f (y1 - y0), y0 = T0 x, y1 = T1 x with T0, T1 the node matrices at `t0` (T1 at
// t0 + S*S): entries in row-major order, every row's two sums over the columns in ascending
// order (the sums of the reference kernel, codegen.py:81-94), in FP64, the row finished as soon
// as its sums are.
template <int S, typename P>
__device__ __forceinline__ void interp_product(P t0, const double (&x)[S], double f,
                                               double (&out)[S]) {
#pragma unroll
    for (int r = 0; r < S; ++r) {
        double y0 = 0.0, y1 = 0.0;
#pragma unroll
        for (int c = 0; c < S; ++c) {
            y0 = fma((double)t0[r * S + c], x[c], y0);
            y1 = fma((double)t0[S * S + r * S + c], x[c], y1);
        }
        out[r] = fma(f, y1 - y0, y0);
    }
}

// One instance per thread, tiles of 256 consecutive instances; workgroup w walks tiles
// w * tiles_per_group ... (w + 1) * tiles_per_group - 1 in order, so the node matrices its
// wavefronts read by scalar loads stay in the scalar cache from one tile to the next (instances
// in segment order: neighbouring tiles sit on neighbouring voltages); the next tile's voltages
// and states are loaded while the current one is computed.
template <int S>
__global__ void __launch_bounds__(256) k_lti_grouped(LtiGroupArgs g) {
    const int nb = g.num_nodes - 1;
    const int64_t tiles = (g.M + 255) / 256;
    int64_t tile = (int64_t)blockIdx.x * g.tiles_per_group;
    const int64_t tile_end = tile + g.tiles_per_group < tiles ? tile + g.tiles_per_group : tiles;
    // the loads of tile `t` into v_n, x_n (FP32 as stored)
    float v_n = 0.0f, x_n[S];
    auto load = [&](int64_t t) {
        const int64_t i = t * 256 + threadIdx.x;
        if (t < tile_end && i < g.M) {
            v_n = g.voltage[g.seg_index ? (int64_t)g.seg_index[i] : i];
#pragma unroll
            for (int k = 0; k < S; ++k) x_n[k] = g.states[k * g.stride + i];
        }
    };
    load(tile);
    for (; tile < tile_end; ++tile) {
        const int64_t i = tile * 256 + threadIdx.x;
        const bool active = i < g.M;
        int b = -1;
        double f = 0.0;
        double x[S];
#pragma unroll
        for (int k = 0; k < S; ++k) x[k] = active ? (double)x_n[k] : 0.0;
        if (active) bucket_of<double>((double)v_n, g.vmin, g.bucket_frq, nb, b, f);
        load(tile + 1);
        // distinct buckets among the wave's active lanes, counted up to group_max + 1
        const uint64_t lanes = __ballot(active);
        int distinct = 0;
        for (uint64_t scan = lanes; scan && distinct <= g.group_max; ++distinct) {
            const int lead = __builtin_ctzll(scan);
            const int u = __builtin_amdgcn_readlane(b, lead);
            scan &= ~__ballot(b == u);
        }
        double out[S];
        if (distinct <= g.group_max) {
            // one pass per distinct bucket: its two node matrices by scalar loads, the lanes of
            // the bucket active
            for (uint64_t rem = lanes; rem;) {
                const int lead = __builtin_ctzll(rem);
                const int u = __builtin_amdgcn_readlane(b, lead);
                const bool mine = b == u;
                rem &= ~__ballot(mine);
                // the address from a copy of u the compiler cannot equate with the lane's own b
                // (inside the branch they are equal, and an address in b would be per lane)
                int ua = u;
                asm volatile("" : "+s"(ua));
                if (mine) {
                    ConstF64 t0 = (ConstF64)(g.f64 + (int64_t)ua * S * S);
                    interp_product<S>(t0, x, f, out);
                }
            }
        } else if (active) {
            // every lane its own node matrices: 16-byte loads of the padded FP32 nodes, consumed
            // in the same row-major order as the grouped path
            const float4* t4 = reinterpret_cast<const float4*>(g.f32_pad + (int64_t)b * g.pad);
            const int q1 = g.pad / 4;
            double y0 = 0.0, y1 = 0.0;
#pragma unroll
            for (int q = 0; q < (S * S + 3) / 4; ++q) {
                const float4 a0 = t4[q], a1 = t4[q1 + q];
                const float e0[4] = {a0.x, a0.y, a0.z, a0.w}, e1[4] = {a1.x, a1.y, a1.z, a1.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int e = 4 * q + j;
                    if (e < S * S) {
                        y0 = fma((double)e0[j], x[e % S], y0);
                        y1 = fma((double)e1[j], x[e % S], y1);
                        if (e % S == S - 1) {
                            out[e / S] = fma(f, y1 - y0, y0);
                            y0 = y1 = 0.0;
                        }
                    }
                }
            }
        }
        if (!active) continue;
        if (g.conserve_sum > 0.0f) {  // codegen.py:95-102
            double tot = 0.0;
#pragma unroll
            for (int r = 0; r < S; ++r) tot += out[r];
            const double corr = (double)g.conserve_sum / tot;
#pragma unroll
            for (int r = 0; r < S; ++r) out[r] *= corr;
        }
#pragma unroll
        for (int k = 0; k < S; ++k) g.states[k * g.stride + i] = (float)out[k];
        if (g.cond_out) {
            double w = 0.0;
#pragma unroll
            for (int k = 0; k < S; ++k) w = fma((double)g.weights[k], (double)(float)out[k], w);
            g.cond_out[i] = (float)((double)g.gmax[i] * w);
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
typedef double f64x4 __attribute__((ext_vector_type(4)));

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
    const size_t n_pad = (size_t)num_nodes * t->pad_stride, n64 = (size_t)num_nodes * S * S;
    hipError_t e = hipMalloc(&t->f32_pad, sizeof(float) * n_pad);
    if (e == hipSuccess) e = hipMalloc(&t->f64, sizeof(double) * n64);
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
    // tiles of 256 instances, a run of consecutive tiles per workgroup, as many workgroups as
    // are resident at once on the device (one round: no tail of late workgroups)
    const int64_t tiles = (M + 255) / 256;
    hipStream_t s = (hipStream_t)stream;
    auto launch = [&](auto kernel) -> int {
        int per_cu = 0;
        NW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
        const int64_t resident = std::max<int64_t>(1, (int64_t)std::max(per_cu, 1) * t->num_cus);
        const int64_t groups = std::min<int64_t>(tiles, resident);
        g.tiles_per_group = (tiles + groups - 1) / groups;
        const unsigned grid = (unsigned)((tiles + g.tiles_per_group - 1) / g.tiles_per_group);
        kernel<<<grid, 256, 0, s>>>(g);
        return NW_OK;
    };
    int rc = NW_OK;
#define NW_LTI_G(K)                                              \
    case K:                                                      \
        rc = launch(k_lti_grouped<K>);                           \
        break;
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
