// Clock events on the device: play back into columns, recording into a ring, and exponentially
// weighted statistics, one launch per clock tick (replaces the Python callbacks of
// neuwon/database/time.py that Clock.tick runs, time.py:114-117, 224-232, 287-302, 509-515,
// 627-637).  The clock's tick count lives in device memory so that a captured step graph replays
// successive ticks; the event table and its length are read from device memory for the same
// reason.  The work per tick is a handful of scalar loads and stores: the launch is latency bound,
// one workgroup, and its cost is the launch itself.
#include "nw_internal.h"

namespace nw {

__device__ __forceinline__ double load_as_double(const void* column, int32_t dtype, int64_t row) {
    switch (dtype) {
        case NW_DTYPE_U8: return (double)static_cast<const uint8_t*>(column)[row];
        case NW_DTYPE_I32: return (double)static_cast<const int32_t*>(column)[row];
        default: return (double)static_cast<const float*>(column)[row];
    }
}

// Trace._callback (time.py:509-515) with each operation rounded on its own.
__device__ void trace_update(double* st, double value, double alpha, double beta) {
#pragma clang fp contract(off)
    double mean = st[0], var = st[1], start = st[2];
    mean = mean + beta * (value - mean);
    start = start * alpha;
    const double true_mean = mean / (1.0 - start);
    const double d = value - true_mean;
    var = var + beta * (d * d);
    var = var * alpha;
    st[0] = mean;
    st[1] = var;
    st[2] = start;
}

__device__ void play(const nw_event& e, int64_t t) {
    if (t < e.first_tick || t >= e.end_tick || e.length <= 0) return;
    int64_t k = e.first_index + (t - e.first_tick);
    if (e.loop) k %= e.length;
    if (k >= e.length) return;
    float* col = static_cast<float*>(e.column);
    const double w = e.wave[k];
    col[e.row] = e.kind == NW_EVENT_SET ? (float)w : (float)(w + (double)col[e.row]);
}

__device__ void sample(const nw_event& e, int64_t t, float* ring, int64_t cap) {
    if (t < e.first_tick || t >= e.end_tick) return;
    const double x = load_as_double(e.column, e.dtype, e.row);
    if (e.kind == NW_EVENT_RECORD)
        ring[e.slot * cap + t % cap] = (float)x;
    else
        trace_update(e.state, x, e.alpha, e.beta);
}

__global__ __launch_bounds__(256) void k_clock_tick(int64_t* tick, const int32_t* num_events,
                                                    const nw_event* ev, float* ring, int64_t cap) {
    const int64_t t = *tick + 1;
    const int n = *num_events;
    int i = 0;
    while (i < n) {
        if (ev[i].kind <= NW_EVENT_ADD) {
            if (threadIdx.x == 0) play(ev[i], t);
            __threadfence_block();
            __syncthreads();
            ++i;
        } else {
            const int end = min(ev[i].run_end, n);
            for (int k = i + (int)threadIdx.x; k < end; k += (int)blockDim.x) sample(ev[k], t, ring, cap);
            __syncthreads();
            i = max(end, i + 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *tick = t;
}

__global__ void k_trace_all(int64_t n, const float* __restrict__ x, float* __restrict__ mean,
                            float* __restrict__ var, double alpha, double beta) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double diff = (double)x[i] - (double)mean[i];
    const double incr = beta * diff;
    mean[i] = (float)((double)mean[i] + incr);
    if (var) var[i] = (float)(alpha * ((double)var[i] + diff * incr));
}

}  // namespace nw

using namespace nw;

extern "C" {

int nw_clock_tick(int64_t* tick, const int32_t* num_events, const nw_event* events, float* ring,
                  int64_t ring_capacity, void* stream) {
    NW_REQUIRE(tick && num_events && events, "nw_clock_tick: null argument");
    NW_REQUIRE(ring_capacity > 0 || ring == nullptr, "nw_clock_tick: ring capacity must be positive");
    k_clock_tick<<<1, 256, 0, (hipStream_t)stream>>>(tick, num_events, events, ring,
                                                     ring_capacity > 0 ? ring_capacity : 1);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_trace_all(int64_t n, const float* x, float* mean, float* var, double alpha, double beta,
                 void* stream) {
    NW_REQUIRE(n >= 0 && (n == 0 || (x && mean)), "nw_trace_all: null argument");
    if (n == 0) return NW_OK;
    k_trace_all<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(n, x, mean, var,
                                                                             alpha, beta);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

}  // extern "C"
