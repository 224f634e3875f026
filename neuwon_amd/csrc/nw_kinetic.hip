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
};

// Waves whose lanes fall into at most this many distinct buckets take the grouped path.
constexpr int kLtiGroupMax = 2;

// y0 = T0 x, y1 = T1 x with T0, T1 the node matrices at `t0` (T1 at t0 + S*S): entries in row-major
// order, so every output row accumulates over the columns in ascending order (the sums of the
// reference kernel, codegen.py:81-94), in FP64.
template <int S, typename P>
__device__ __forceinline__ void node_products(P t0, const double (&x)[S], double (&y0)[S],
                                              double (&y1)[S]) {
#pragma unroll
    for (int r = 0; r < S; ++r) y0[r] = y1[r] = 0.0;
#pragma unroll
    for (int e = 0; e < S * S; ++e) {
        y0[e / S] = fma((double)t0[e], x[e % S], y0[e / S]);
        y1[e / S] = fma((double)t0[S * S + e], x[e % S], y1[e / S]);
    }
}

template <int S>
__global__ void __launch_bounds__(256) k_lti_grouped(LtiGroupArgs g) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool active = i < g.M;
    const int nb = g.num_nodes - 1;
    int b = -1;
    double f = 0.0;
    double x[S];
#pragma unroll
    for (int k = 0; k < S; ++k) x[k] = 0.0;
    if (active) {
        const double v = g.voltage[g.seg_index ? (int64_t)g.seg_index[i] : i];
        bucket_of<double>(v, g.vmin, g.bucket_frq, nb, b, f);
#pragma unroll
        for (int k = 0; k < S; ++k) x[k] = g.states[k * g.stride + i];
    }
    // distinct buckets among the wave's active lanes, counted up to group_max + 1
    const uint64_t lanes = __ballot(active);
    int distinct = 0;
    for (uint64_t scan = lanes; scan && distinct <= g.group_max; ++distinct) {
        const int lead = __builtin_ctzll(scan);
        const int u = __builtin_amdgcn_readlane(b, lead);
        scan &= ~__ballot(b == u);
    }
    double y0[S], y1[S];
    if (distinct <= g.group_max) {
        // one pass per distinct bucket: its two node matrices by scalar loads, the lanes of the
        // bucket active
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
                double z0[S], z1[S];
                node_products<S>(t0, x, z0, z1);
#pragma unroll
                for (int r = 0; r < S; ++r) {
                    y0[r] = z0[r];
                    y1[r] = z1[r];
                }
            }
        }
    } else if (active) {
        // every lane its own node matrices: 16-byte loads of the padded FP32 nodes, consumed in
        // the same row-major order as the grouped path
        const float4* t4 = reinterpret_cast<const float4*>(g.f32_pad + (int64_t)b * g.pad);
        const int q1 = g.pad / 4;
#pragma unroll
        for (int r = 0; r < S; ++r) y0[r] = y1[r] = 0.0;
#pragma unroll
        for (int q = 0; q < (S * S + 3) / 4; ++q) {
            const float4 a0 = t4[q], a1 = t4[q1 + q];
            const float e0[4] = {a0.x, a0.y, a0.z, a0.w}, e1[4] = {a1.x, a1.y, a1.z, a1.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int e = 4 * q + j;
                if (e < S * S) {
                    y0[e / S] = fma((double)e0[j], x[e % S], y0[e / S]);
                    y1[e / S] = fma((double)e1[j], x[e % S], y1[e / S]);
                }
            }
        }
    }
    if (!active) return;
    double out[S];
#pragma unroll
    for (int r = 0; r < S; ++r) out[r] = fma(f, y1[r] - y0[r], y0[r]);
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

}  // namespace nw

using namespace nw;

extern "C" {

int nw_lti_table_create(int device, int32_t S, int32_t num_nodes, float vmin, float vmax,
                        const float* node_table, nw_lti_table** out) {
    NW_REQUIRE(out && node_table, "nw_lti_table_create: null argument");
    NW_REQUIRE(S >= 1 && S <= 16, "state count %d not supported (1..16)", S);
    NW_REQUIRE(num_nodes >= 2 && vmax > vmin, "bad table grid");
    *out = nullptr;
    NW_HIP(hipSetDevice(device));
    nw_lti_table* t = new nw_lti_table;
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
    const unsigned grid = (unsigned)((M + 255) / 256);
    hipStream_t s = (hipStream_t)stream;
#define NW_LTI_G(K)                                              \
    case K:                                                      \
        k_lti_grouped<K><<<grid, 256, 0, s>>>(g);                \
        break;
    switch (t->S) {
        NW_LTI_G(1) NW_LTI_G(2) NW_LTI_G(3) NW_LTI_G(4) NW_LTI_G(5) NW_LTI_G(6) NW_LTI_G(7)
        NW_LTI_G(8) NW_LTI_G(9) NW_LTI_G(10) NW_LTI_G(11) NW_LTI_G(12) NW_LTI_G(13)
        NW_LTI_G(14) NW_LTI_G(15) NW_LTI_G(16)
        default: NW_REQUIRE(false, "state count %d not supported (1..16)", t->S);
    }
#undef NW_LTI_G
    NW_HIP(hipGetLastError());
    return NW_OK;
}

}  // extern "C"
