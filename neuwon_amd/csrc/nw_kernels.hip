// Fused staggered step for HH cells, the LTI kinetic kernel and the segmented accumulation.
//
// Fused step (replaces RxD_Model.advance, neuwon/rxd/rxd_model.py:63-123, for cells whose every
// segment carries leak + HH Na + HH K written as two-state kinetic schemes).  Cells are
// independent (the propagator is block diagonal), so one workgroup advances one cell through any
// number of steps without a grid-wide barrier.  Per step:
//   A  x = relax(v, g, d, I)                 thread per segment, x -> LDS      (electric.py:71-74)
//   B  v = P x                               wavefront per CSR row, v -> LDS   (electric.py:75)
//   C  gates <- LTI table(v); g, d; x        thread per segment                (codegen.py:39-106,
//                                                                               species.py:155-160)
//   D  v = P x                               wavefront per CSR row
//   AP detector and optional trace of one segment per cell                     (neuron.py:72-76)
// The voltage of the cell stays in LDS across steps; gates and per-segment constants stream from
// HBM (they are a few percent of the propagator's bytes).
#include <algorithm>
#include <mutex>
#include <cstring>

#include <type_traits>

#include "nw_internal.h"

namespace nw {

struct HHArgs {
    const int64_t* cell_row0;
    const int64_t* cell_nz0;
    const uint32_t* row_off;
    const float* vals;
    const uint16_t* cols;
    const double* rowsum;
    const int32_t* order;
    int32_t nmax;
    int32_t n_steps;
    nw_hh_params prm;
    const float* table;  // [3][num_nodes][2]
    float* v;
    float* m;
    float* h;
    float* n;
    const float* g_na;
    const float* g_k;
    const float* g_l;
    const float* inv_c;
    const int32_t* stim_local;
    const int32_t* ap_local;
    uint8_t* ap_state;
    uint32_t* ap_count;
    float* trace;
    int64_t trace_stride;
    float* record;       // optional: per step 5 arrays (v, m, h, n at the start, v after half A)
    int64_t num_rows;
    const double* wgt;   // symmetric format: w of every segment
    ChebStore cheb;      // Chebyshev store
    const int64_t* step_dev;  // graph replays: index of the first step in device memory (else
                              // prm.step0)
};

// Stimulus current (A) of the cell's stimulated segment on input tick `tick` (1-based).
__device__ __forceinline__ float stim_current(const nw_hh_params& p, uint32_t cell, int64_t tick) {
    if (p.stim_mode == 1) return p.stim_amp_a;
    if (p.stim_mode == 2) {
        int k = 0;
        for (int j = 0; j < p.pulse_ticks; ++j) {
            int64_t t = tick - j;
            if (t < 0) break;
            k += stim_hash(p.stim_seed, cell, (uint32_t)t) < p.poisson_threshold ? 1 : 0;
        }
        return p.stim_amp_a * (float)k;
    }
    return 0.0f;
}

// open' = a(v) + b(v) * open with (a, b) linearly interpolated between grid nodes.
__device__ __forceinline__ double gate_advance(const float2* tbl, int b, double f, double open) {
    const float2 t0 = tbl[b];
    const float2 t1 = tbl[b + 1];
    const double a = fma(f, (double)t1.x - (double)t0.x, (double)t0.x);
    const double s = fma(f, (double)t1.y - (double)t0.y, (double)t0.y);
    return fma(s, open, a);
}

// hh.mod:88-94; x / (exp(x/y) - 1) evaluated with expm1 (same value, no cancellation near 0).
__device__ __forceinline__ double vtrap(double x, double y) {
    const double r = x / y;
    return (fabs(r) < 1e-6) ? y * (1.0 - r * 0.5) : x / expm1(r);
}

// Exact update of one hh.mod gate (solver.py:48-55): x' = inf + (x - inf) exp(-(a+b) q10 dt).
__device__ __forceinline__ double gate_exact(double alpha, double beta, double q10, double dt,
                                             double x) {
    const double s = q10 * (alpha + beta);
    const double inf = alpha / (alpha + beta);
    return fma(x - inf, exp(-s * dt), inf);
}

__device__ __forceinline__ void hh_exact(double v, double q10, double dt, double& m, double& h,
                                         double& n) {
    // rates, hh.mod:62-86
    const double am = 0.1 * vtrap(-(v + 40.0), 10.0);
    const double bm = 4.0 * exp(-(v + 65.0) / 18.0);
    const double ah = 0.07 * exp(-(v + 65.0) / 20.0);
    const double bh = 1.0 / (exp(-(v + 35.0) / 10.0) + 1.0);
    const double an = 0.01 * vtrap(-(v + 55.0), 10.0);
    const double bn = 0.125 * exp(-(v + 65.0) / 80.0);
    m = gate_exact(am, bm, q10, dt, m);
    h = gate_exact(ah, bh, q10, dt, h);
    n = gate_exact(an, bn, q10, dt, n);
}

// Breakpoint of the three channels (hh.mod:41-46) and the driving voltage sum(g E)/sum(g)
// (rxd_model.py:105-109; 0 where the conductance is 0).
__device__ __forceinline__ void hh_conductance(double gna_bar, double gk_bar, double gl, double m,
                                               double h, double n, double e_na, double e_k,
                                               double e_l, double& g, double& d) {
    const double gna = gna_bar * m * m * m * h;
    const double gk = gk_bar * n * n * n * n;
    g = gna + gk + gl;
    const double num = gna * e_na + gk * e_k + gl * e_l;
    d = (g != 0.0) ? num / g : 0.0;
}

// Every per-segment quantity is computed in FP64 registers from FP32 storage; voltages live in LDS
// as FP32 offsets from the cell's reference voltage and the row products accumulate in FP64.
//
// SYM = false: general CSR (spmv_rows); LDS holds x - vref and v - vref (FP32).
// SYM = true:  symmetric CSR of S = W P W^-1 (spmv_sym_rows); LDS holds y = W (x - vref) (FP32) and
//              z = S y (FP64), from which v_i = vref * rowsum_i + z_i / w_i.  z is cleared by the
//              thread that reads it, before the next product accumulates into it.
template <int BLOCK, int GATE_MODE, bool SYM>
__global__ void __launch_bounds__(BLOCK) k_hh_step(HHArgs a) {
    extern __shared__ double lds64[];
    float* xs;                  // x - vref (general) or y (symmetric) of the current half step
    float* vs = nullptr;        // v - vref (general)
    double* zs = nullptr;       // z (symmetric)
    if (SYM) {
        zs = lds64;
        xs = reinterpret_cast<float*>(lds64 + a.nmax);
    } else {
        xs = reinterpret_cast<float*>(lds64);
        vs = xs + a.nmax;
    }
    const nw_hh_params& p = a.prm;
    const int cell = a.order[blockIdx.x];
    const int64_t row0 = a.cell_row0[cell];
    const int nrows = (int)(a.cell_row0[cell + 1] - row0);
    const uint32_t* roff = a.row_off + row0 + cell;
    const float* vals = a.vals + a.cell_nz0[cell];
    const uint16_t* cols = a.cols + a.cell_nz0[cell];
    const double* rs = a.rowsum + row0;
    const double* wg = SYM ? a.wgt + row0 : nullptr;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    constexpr int NW = BLOCK / kWave;
    const double hdt = 0.5 * (double)p.dt_ms;
    const double hdt_s = hdt * 1e-3;
    const int nb = p.num_nodes - 1;
    const double vmin = p.vmin;
    const double bucket_frq = (double)nb / ((double)p.vmax - (double)p.vmin);
    const double e_na = p.e_na, e_k = p.e_k, e_l = p.e_l;
    const float2* tm = reinterpret_cast<const float2*>(a.table);
    const float2* th = tm + p.num_nodes;
    const float2* tn = th + p.num_nodes;
    const int stim_row = a.stim_local ? a.stim_local[cell] : -1;
    const int ap_row = a.ap_local ? a.ap_local[cell] : -1;
    uint8_t ap_state = 0;
    uint32_t ap_count = 0;
    if (threadIdx.x == 0 && ap_row >= 0) {
        ap_state = a.ap_state[cell];
        ap_count = a.ap_count[cell];
    }

    const double vref = (double)a.v[row0];  // the cell's reference voltage for this launch
    // voltage of segment i after a product; in the symmetric form the thread that owns i clears z_i
    auto take_v = [&](int i) -> double {
        if (SYM) {
            const double z = zs[i];
            zs[i] = 0.0;
            return fma(vref, rs[i], z / wg[i]);
        }
        return vref + (double)vs[i];
    };
    auto peek_v = [&](int i) -> double {
        return SYM ? fma(vref, rs[i], zs[i] / wg[i]) : vref + (double)vs[i];
    };
    auto put_x = [&](int i, double x) {
        xs[i] = SYM ? (float)((x - vref) * wg[i]) : (float)(x - vref);
    };
    auto product = [&]() {
        if (SYM) spmv_sym_rows(roff, vals, cols, xs, zs, nrows, wave, NW, lane);
        else spmv_rows(roff, vals, cols, rs, xs, vs, vref, nrows, wave, NW, lane);
    };
    for (int i = threadIdx.x; i < nrows; i += BLOCK) {
        const double v = (double)a.v[row0 + i];
        if (SYM) zs[i] = (v - vref * rs[i]) * wg[i];
        else vs[i] = (float)(v - vref);
    }
    __syncthreads();

    for (int step = 0; step < a.n_steps; ++step) {
        const int64_t gstep = p.step0 + step;
        const int64_t tick_a = 2 * gstep + 1;
        const bool g_valid = p.conductances_valid || step > 0;
        // ---- A: relax with the conductances of the previous mechanism step
        for (int i = threadIdx.x; i < nrows; i += BLOCK) {
            const int64_t gi = row0 + i;
            const double invc = a.inv_c[gi];
            const double v = take_v(i);
            double g = 0.0, d = 0.0;
            if (g_valid)
                hh_conductance(a.g_na[gi], a.g_k[gi], a.g_l[gi], a.m[gi], a.h[gi], a.n[gi], e_na,
                               e_k, e_l, g, d);
            const double I = (i == stim_row) ? (double)stim_current(p, (uint32_t)cell, tick_a) : 0.0;
            const double alpha = exp(-hdt_s * g * invc);
            put_x(i, d - (d - v) * alpha + hdt * I * invc);
            if (a.record) {
                float* rec = a.record + (int64_t)step * 5 * a.num_rows + gi;
                rec[0] = (float)v;
                rec[a.num_rows] = a.m[gi];
                rec[2 * a.num_rows] = a.h[gi];
                rec[3 * a.num_rows] = a.n[gi];
            }
        }
        __syncthreads();
        // ---- B: v = P x
        product();
        __syncthreads();
        // ---- C: mechanisms at v(t + dt/2), then relax for the second half step
        for (int i = threadIdx.x; i < nrows; i += BLOCK) {
            const int64_t gi = row0 + i;
            const double v = take_v(i);
            if (a.record) a.record[((int64_t)step * 5 + 4) * a.num_rows + gi] = (float)v;
            double m = a.m[gi], h = a.h[gi], n = a.n[gi];
            if (GATE_MODE == 1) {
                hh_exact(v, (double)p.q10, (double)p.dt_ms, m, h, n);
            } else {
                int b;
                double f;
                bucket_of(v, vmin, bucket_frq, nb, b, f);
                m = gate_advance(tm, b, f, m);
                h = gate_advance(th, b, f, h);
                n = gate_advance(tn, b, f, n);
            }
            const float m32 = (float)m, h32 = (float)h, n32 = (float)n;
            a.m[gi] = m32;
            a.h[gi] = h32;
            a.n[gi] = n32;
            double g, d;  // conductances of the stored (FP32) states
            hh_conductance(a.g_na[gi], a.g_k[gi], a.g_l[gi], m32, h32, n32, e_na, e_k, e_l, g, d);
            const double invc = a.inv_c[gi];
            const double I =
                (i == stim_row) ? (double)stim_current(p, (uint32_t)cell, tick_a + 1) : 0.0;
            const double alpha = exp(-hdt_s * g * invc);
            put_x(i, d - (d - v) * alpha + hdt * I * invc);
        }
        __syncthreads();
        // ---- D: v = P x
        product();
        __syncthreads();
        if (ap_row >= 0) {
            if (threadIdx.x == 0) {
                const double vap = peek_v(ap_row);
                const bool over = vap >= (double)p.ap_threshold;
                if (over && !ap_state) ++ap_count;
                ap_state = over ? 1 : 0;
                if (a.trace) a.trace[(int64_t)cell * a.trace_stride + step] = (float)vap;
            }
            if (SYM) __syncthreads();  // z[ap_row] is cleared in the next phase A
        }
    }
    for (int i = threadIdx.x; i < nrows; i += BLOCK) {
        const float vf = (float)peek_v(i);
        a.v[row0 + i] = vf;
        if (a.record) {  // the state after the last step closes the record
            float* rec = a.record + (int64_t)a.n_steps * 5 * a.num_rows + row0 + i;
            rec[0] = vf;
            rec[a.num_rows] = a.m[row0 + i];
            rec[2 * a.num_rows] = a.h[row0 + i];
            rec[3 * a.num_rows] = a.n[row0 + i];
        }
    }
    if (threadIdx.x == 0 && ap_row >= 0) {
        a.ap_state[cell] = ap_state;
        a.ap_count[cell] = ap_count;
    }
}

// ----------------------------------------------------------------------------------------------
// The step on the Chebyshev store, as two half steps (HALF 0: relax with the conductances of the
// previous mechanism step, then v = P x; HALF 1: gates at v(t + dt/2), relax, v = P x, AP
// detector), run one after the other by a cell's workgroup (k_hh_step_cheb: one launch per step;
// k_hh_step2_cheb: two steps per launch) or as one launch each (k_hh_half_cheb); the launch form
// is nw_hh_params.cheb_launch.  The series are sequential code regions, which keeps a launch
// within the 128 registers of 1024-thread blocks.  Voltages pass between launches as FP64 (v64)
// and within a launch in LDS; the FP32 voltage array is written after a launch's last step.
// ----------------------------------------------------------------------------------------------
// Gate table nodes the second half step's kernel keeps in LDS (3 gates x 2 FP32 x 2048 nodes,
// 48 KB beside the 88 KB of the five-row-group term vectors); larger tables are read from memory.
constexpr int kGateNodesLds = 2048;
// Whether the second half step of a launch shape keeps the gate tables in LDS.
template <int GATE_MODE, int MAXJ, int KIND>
constexpr bool gate_tables_in_lds() {
    return GATE_MODE == 0 && MAXJ > 1 &&
           (prod_lds(MAXJ, KIND) + 3 * kGateNodesLds) * 8 <= 160 * 1024;
}

// One half step of one cell (the body of k_hh_half_cheb / k_hh_step_cheb): b0 the product's LDS,
// gtab the LDS of the gate tables (gate_tables_in_lds).
// FUSED: the two halves run in one workgroup (k_hh_step_cheb, k_hh_step2_cheb): the first leaves
// v(t + dt/2) in LDS at the rows' places, where the second reads it, instead of passing it through
// v64.  RUN (k_hh_step2_cheb, two steps in one workgroup): the second half leaves v(t + dt) in
// LDS for the next step's first half (v_lds), and the voltage arrays are written after the
// launch's last step only (last).
template <int GATE_MODE, int HALF, int MAXJ, int NT, int KIND, bool FUSED = false, bool RUN = false>
__device__ __forceinline__ void hh_half_cheb(const HHArgs& a, double* v64, int step, int from_v32,
                                             double* b0, float2* gtab, bool v_lds = false,
                                             bool last = true) {
    const nw_hh_params& p = a.prm;
    const int cell = a.order[blockIdx.x];
    const int64_t row0 = a.cell_row0[cell];
    const int nrows = (int)(a.cell_row0[cell + 1] - row0);
    const int64_t pre_base = a.cheb.pre_base[cell];
    const bool pre = pre_base >= 0;
    const uint16_t* pos = a.cheb.pre_pos + row0;
    const double hdt = 0.5 * (double)p.dt_ms;
    const double hdt_s = hdt * 1e-3;
    const int64_t gstep = (a.step_dev ? *a.step_dev : p.step0) + step;
    const int64_t tick = 2 * gstep + 1 + HALF;
    const int stim_row = a.stim_local ? a.stim_local[cell] : -1;
    const bool g_valid = HALF == 1 || p.conductances_valid || step > 0;
    const int64_t NR = a.num_rows;
    // The thread's rows i = threadIdx.x + j * 1024 (j < MAXJ) in groups of at most three, each in
    // phases so that the loads of the group are in flight together: the per-segment values, then
    // (table gates) the gate table entries at each row's voltage, then the arithmetic.
    constexpr int R = MAXJ;
    constexpr int G = R < 3 ? R : 3;
    // Gate tables in LDS (one workgroup per CU: the LDS beside the term vectors is free): the
    // copy's loads are issued with the first group's, so each row's lookups are LDS reads instead
    // of a second memory round trip per group.
    constexpr bool kStage = HALF == 1 && gate_tables_in_lds<GATE_MODE, MAXJ, KIND>();
    constexpr int kStageLoads = (3 * kGateNodesLds + NT - 1) / NT;
    const bool stage = kStage && p.num_nodes <= kGateNodesLds;
    float2 tcopy[kStage ? kStageLoads : 1];
    if (stage) {
        const float2* tsrc = reinterpret_cast<const float2*>(a.table);
#pragma unroll
        for (int q = 0; q < kStageLoads; ++q) {
            const int e = threadIdx.x + q * NT;
            tcopy[q] = e < 3 * p.num_nodes ? tsrc[e] : make_float2(0.f, 0.f);
        }
    }
#pragma unroll
    for (int j0 = 0; j0 < R; j0 += G) {
        double v[G];
        float fm[G], fh[G], fn[G], fgna[G], fgk[G], fgl[G], fic[G];
        int place[G];
#pragma unroll
        for (int jj = 0; jj < G; ++jj) {
            const int j = j0 + jj;
            const int i = threadIdx.x + j * kChebBlock;
            if (j < R && i < nrows) {
                const int64_t gi = row0 + i;
                if ((FUSED && HALF == 1) || (RUN && v_lds))
                    v[jj] = b0[pre ? (int)pos[i] : i];
                else
                    v[jj] = from_v32 ? (double)a.v[gi] : v64[gi];
                fm[jj] = a.m[gi];
                fh[jj] = a.h[gi];
                fn[jj] = a.n[gi];
                fgna[jj] = a.g_na[gi];
                fgk[jj] = a.g_k[gi];
                fgl[jj] = a.g_l[gi];
                fic[jj] = a.inv_c[gi];
                place[jj] = pre ? (int)pos[i] : i;
            }
        }
        float2 tab[G][3][2];
        double frac[G];
        if (HALF == 1 && GATE_MODE == 0) {
            if (kStage && j0 == 0 && stage) {
#pragma unroll
                for (int q = 0; q < kStageLoads; ++q) {
                    const int e = threadIdx.x + q * NT;
                    if (e < 3 * kGateNodesLds) gtab[e] = tcopy[q];
                }
            }
            if (kStage && j0 == 0) __syncthreads();  // (stage is uniform over the workgroup)
            const int nb = p.num_nodes - 1;
            const double frq = (double)nb / ((double)p.vmax - (double)p.vmin);
            const float2* tm = reinterpret_cast<const float2*>(a.table);
#pragma unroll
            for (int jj = 0; jj < G; ++jj) {
                const int j = j0 + jj;
                const int i = threadIdx.x + j * kChebBlock;
                if (j < R && i < nrows) {
                    int bk;
                    bucket_of(v[jj], (double)p.vmin, frq, nb, bk, frac[jj]);
                    if (stage) {
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            tab[jj][q][0] = gtab[q * p.num_nodes + bk];
                            tab[jj][q][1] = gtab[q * p.num_nodes + bk + 1];
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            tab[jj][q][0] = tm[q * p.num_nodes + bk];
                            tab[jj][q][1] = tm[q * p.num_nodes + bk + 1];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int jj = 0; jj < G; ++jj) {
            const int j = j0 + jj;
            const int i = threadIdx.x + j * kChebBlock;
            if (j >= R || i >= nrows) continue;
            const int64_t gi = row0 + i;
            double g = 0.0, d = 0.0;
            if (a.record) {  // the record of k_hh_step: v, m, h, n at the start, v after half A
                float* rec = a.record + (int64_t)step * 5 * NR + gi;
                if (HALF == 0) {
                    rec[0] = (float)v[jj];
                    rec[NR] = fm[jj];
                    rec[2 * NR] = fh[jj];
                    rec[3 * NR] = fn[jj];
                } else {
                    rec[4 * NR] = (float)v[jj];
                }
            }
            if (HALF == 1) {
                double m = fm[jj], h = fh[jj], n = fn[jj];
                if (GATE_MODE == 1) {
                    hh_exact(v[jj], (double)p.q10, (double)p.dt_ms, m, h, n);
                } else {  // open' = a + b open, (a, b) interpolated between the two table nodes
                    double* x3[3] = {&m, &h, &n};
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const float2 t0 = tab[jj][q][0], t1 = tab[jj][q][1];
                        const double f = frac[jj];
                        const double ga = fma(f, (double)t1.x - (double)t0.x, (double)t0.x);
                        const double gb = fma(f, (double)t1.y - (double)t0.y, (double)t0.y);
                        *x3[q] = fma(gb, *x3[q], ga);
                    }
                }
                fm[jj] = (float)m;
                fh[jj] = (float)h;
                fn[jj] = (float)n;
                a.m[gi] = fm[jj];
                a.h[gi] = fh[jj];
                a.n[gi] = fn[jj];
                hh_conductance(fgna[jj], fgk[jj], fgl[jj], fm[jj], fh[jj], fn[jj], p.e_na, p.e_k,
                               p.e_l, g, d);
            } else if (g_valid) {
                hh_conductance(fgna[jj], fgk[jj], fgl[jj], fm[jj], fh[jj], fn[jj], p.e_na, p.e_k,
                               p.e_l, g, d);
            }
            const double invc = fic[jj];
            const double I = (i == stim_row) ? (double)stim_current(p, (uint32_t)cell, tick) : 0.0;
            const double alpha = exp(-hdt_s * g * invc);
            b0[place[jj]] = d - (d - v[jj]) * alpha + hdt * I * invc;
        }
    }
    __syncthreads();
    double out[kChebRows];
    const int places = cheb_cell<MAXJ, NT, KIND>(a.cheb, row0, cell, nrows, pre_base, b0, out);
    // results to LDS (the row's place), then out in natural order
#pragma unroll
    for (int j = 0; j < kChebRows; ++j) {
        const int r = threadIdx.x + j * kChebBlock;
        if (r < places) b0[r] = out[j];
    }
    __syncthreads();
    const bool closes = HALF == 1 && a.record && step == a.n_steps - 1;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = threadIdx.x + j * kChebBlock;
        if (i >= nrows) continue;
        const int64_t gi = row0 + i;
        const double vi = b0[pre ? (int)pos[i] : i];
        if (!(FUSED && HALF == 0) && (!RUN || last)) v64[gi] = vi;
        if (HALF == 1 && (!RUN || last)) a.v[gi] = (float)vi;
        if (closes) {  // the state after the last step closes the record
            float* rec = a.record + (int64_t)a.n_steps * 5 * NR + gi;
            rec[0] = (float)vi;
            rec[NR] = a.m[gi];
            rec[2 * NR] = a.h[gi];
            rec[3 * NR] = a.n[gi];
        }
    }
    if (HALF == 1) {
        const int ap_row = a.ap_local ? a.ap_local[cell] : -1;
        if (ap_row >= 0) {
            if (threadIdx.x == 0) {
                const double vap = b0[pre ? pos[ap_row] : ap_row];
                const bool over = vap >= (double)p.ap_threshold;
                if (over && !a.ap_state[cell]) ++a.ap_count[cell];
                a.ap_state[cell] = over ? 1 : 0;
                if (a.trace) a.trace[(int64_t)cell * a.trace_stride + step] = (float)vap;
            }
        }
    }
}

template <int GATE_MODE, int HALF, int MAXJ, int NT = kChebBlock, int KIND = kProdAny>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MAXJ == 1 ? 8 : 4)))
k_hh_half_cheb(HHArgs a, double* v64, int step, int from_v32) {
    __shared__ __align__(16) double b0[prod_lds(MAXJ, KIND)];
    constexpr bool kTab = HALF == 1 && gate_tables_in_lds<GATE_MODE, MAXJ, KIND>();
    __shared__ __align__(16) float2 gtab[kTab ? 3 * kGateNodesLds : 1];
    hh_half_cheb<GATE_MODE, HALF, MAXJ, NT, KIND>(a, v64, step, from_v32, b0, gtab);
}

// Both half steps of one step in one launch: the workgroup of a cell runs its first half step,
// then its second (every cell's two halves depend on that cell alone), so the second half step's
// loads of the cell's segments and store find them in the L2, v(t + dt/2) stays in LDS, and a
// step costs one launch.
template <int GATE_MODE, int MAXJ, int NT = kChebBlock, int KIND = kProdAny>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MAXJ == 1 ? 8 : 4)))
k_hh_step_cheb(HHArgs a, double* v64, int step, int from_v32) {
    __shared__ __align__(16) double b0[prod_lds(MAXJ, KIND)];
    constexpr bool kTab = gate_tables_in_lds<GATE_MODE, MAXJ, KIND>();
    __shared__ __align__(16) float2 gtab[kTab ? 3 * kGateNodesLds : 1];
    hh_half_cheb<GATE_MODE, 0, MAXJ, NT, KIND, true>(a, v64, step, from_v32, b0, gtab);
    __syncthreads();
    hh_half_cheb<GATE_MODE, 1, MAXJ, NT, KIND, true>(a, v64, step, 0, b0, gtab);
}

// Two steps of every cell in one launch: four half steps in sequence, v(t) in LDS between them.
// (A loop over all of a call's steps in one launch keeps what the compiler hoists out of the loop
// live across both series: 67-88 VGPRs spilled; two steps written out spill 12.)
template <int GATE_MODE, int MAXJ, int NT = kChebBlock, int KIND = kProdAny>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MAXJ == 1 ? 8 : 4)))
k_hh_step2_cheb(HHArgs a, double* v64, int step, int from_v32) {
    __shared__ __align__(16) double b0[prod_lds(MAXJ, KIND)];
    constexpr bool kTab = gate_tables_in_lds<GATE_MODE, MAXJ, KIND>();
    __shared__ __align__(16) float2 gtab[kTab ? 3 * kGateNodesLds : 1];
    hh_half_cheb<GATE_MODE, 0, MAXJ, NT, KIND, true, true>(a, v64, step, from_v32, b0, gtab, false, false);
    __syncthreads();
    hh_half_cheb<GATE_MODE, 1, MAXJ, NT, KIND, true, true>(a, v64, step, 0, b0, gtab, true, false);
    __syncthreads();
    hh_half_cheb<GATE_MODE, 0, MAXJ, NT, KIND, true, true>(a, v64, step + 1, 0, b0, gtab, true, true);
    __syncthreads();
    hh_half_cheb<GATE_MODE, 1, MAXJ, NT, KIND, true, true>(a, v64, step + 1, 0, b0, gtab, true, true);
}

// ----------------------------------------------------------------------------------------------
// Reverse pass of n_steps fused steps (table gates), from a record written by k_hh_step.
// Loss: sum over steps of (v[soma] after the step - target)^2.  Adjoints of the voltage and the
// three gates are carried across launches; parameter gradients accumulate in FP64.
//   general store:   P^T y = C P (C^-1 y): a forward product of y/C scaled by C;
//   symmetric store: P^T y = W S (W^-1 y) with the stored S = W P W^-1.
// ----------------------------------------------------------------------------------------------
struct AdjArgs {
    const int64_t* cell_row0;
    const int64_t* cell_nz0;
    const uint32_t* row_off;
    const float* vals;
    const uint16_t* cols;
    const double* rowsum;
    const int32_t* order;
    int32_t nmax;
    int32_t n_steps;
    nw_hh_params prm;
    const float* record;
    int64_t num_rows;
    const float* table6;   // [3][num_nodes][6]: a, b, da/dta, db/dta, da/dtb, db/dtb
    const float* g_na;
    const float* g_k;
    const float* g_l;
    const float* inv_c;
    const int32_t* soma_local;
    const float* target;
    int64_t target_stride;
    float* lam_v;
    float* lam_m;
    float* lam_h;
    float* lam_n;
    double* grad_na;
    double* grad_k;
    double* grad_l;
    double* grad_theta;    // [cell][6]
    double* loss;          // [cell]
    const double* wgt;     // symmetric store: w of every segment
};

struct GateLookup {
    double val[6];
    double slope[6];
};

__device__ __forceinline__ void lookup6(const float* tbl, int num_nodes, double v, double vmin,
                                        double frq, GateLookup& o) {
    const int nb = num_nodes - 1;
    const double u = (v - vmin) * frq;
    int b = (int)u;
    double f;
    bool inside = true;
    if (b > nb - 1) { b = nb - 1; f = 1.0; inside = false; }
    else if (b < 0) { b = 0; f = 0.0; inside = false; }
    else f = u - (double)b;
    const float* t0 = tbl + (int64_t)b * 6;
    const float* t1 = t0 + 6;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double d = (double)t1[q] - (double)t0[q];
        o.val[q] = fma(f, d, (double)t0[q]);
        o.slope[q] = inside ? d * frq : 0.0;
    }
}

// Adjoint of x = D(1-al) + v al + b through G, N of the given gates; returns lambda_v and adds the
// gate adjoints and the conductance gradients.
__device__ __forceinline__ double half_adjoint(double lx, double v_in, double m, double h, double n,
                                               double gna_bar, double gk_bar, double gl,
                                               double e_na, double e_k, double e_l, double hs,
                                               double invc, double& lm, double& lh, double& ln,
                                               double& gr_na, double& gr_k, double& gr_l) {
    const double m3h = m * m * m * h, n4 = n * n * n * n;
    const double gna = gna_bar * m3h, gk = gk_bar * n4;
    const double G = gna + gk + gl;
    const double N = gna * e_na + gk * e_k + gl * e_l;
    const double D = N / G;
    const double al = exp(-hs * G * invc);
    const double lD = (1.0 - al) * lx;
    const double lal = (v_in - D) * lx;
    const double lG = -lal * al * hs * invc - lD * D / G;
    const double lN = lD / G;
    const double l_gna = lG + e_na * lN, l_gk = lG + e_k * lN, l_gl = lG + e_l * lN;
    gr_na += l_gna * m3h;
    gr_k += l_gk * n4;
    gr_l += l_gl;
    lm += l_gna * gna_bar * 3.0 * m * m * h;
    lh += l_gna * gna_bar * m * m * m;
    ln += l_gk * gk_bar * 4.0 * n * n * n;
    return al * lx;
}

template <int BLOCK, bool SYM>
__global__ void __launch_bounds__(BLOCK) k_hh_adjoint(AdjArgs a) {
    extern __shared__ double lds64[];
    // general: ys = adjoint / C and the FP32 product zs32; symmetric: ys = adjoint / w and the FP64
    // product zs (cleared by the thread that reads it); lv: adjoint of the voltage
    float *ys, *zs32 = nullptr, *lv;
    double* zs = nullptr;
    if (SYM) {
        zs = lds64;
        ys = reinterpret_cast<float*>(lds64 + a.nmax);
        lv = ys + a.nmax;
    } else {
        ys = reinterpret_cast<float*>(lds64);
        zs32 = ys + a.nmax;
        lv = zs32 + a.nmax;
    }
    __shared__ double red[6][BLOCK / kWave];
    __shared__ double red_loss[BLOCK / kWave];
    const nw_hh_params& p = a.prm;
    const int cell = a.order[blockIdx.x];
    const int64_t row0 = a.cell_row0[cell];
    const int nrows = (int)(a.cell_row0[cell + 1] - row0);
    const uint32_t* roff = a.row_off + row0 + cell;
    const float* vals = a.vals + a.cell_nz0[cell];
    const uint16_t* cols = a.cols + a.cell_nz0[cell];
    const double* rs = a.rowsum + row0;
    const double* wg = SYM ? a.wgt + row0 : nullptr;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t NR = a.num_rows;
    const double hs = 0.5 * (double)p.dt_ms * 1e-3;
    const double vmin = p.vmin;
    const double frq = (double)(p.num_nodes - 1) / ((double)p.vmax - (double)p.vmin);
    const double e_na = p.e_na, e_k = p.e_k, e_l = p.e_l;
    const int soma = a.soma_local[cell];
    double th[6] = {0, 0, 0, 0, 0, 0};
    double loss = 0.0;
    auto put_y = [&](int i, double lam, double invc) {
        ys[i] = SYM ? (float)(lam / wg[i]) : (float)(lam * invc);
    };
    auto take_lx = [&](int i, double invc) -> double {  // (P^T lambda)_i
        if (SYM) {
            const double z = zs[i];
            zs[i] = 0.0;
            return z * wg[i];
        }
        return (double)zs32[i] / invc;
    };
    auto product = [&]() {
        if (SYM)
            spmv_sym_rows(roff, vals, cols, ys, zs, nrows, __builtin_amdgcn_readfirstlane(wave),
                          BLOCK / kWave, lane);
        else
            spmv<0>(roff, vals, cols, rs, ys, zs32, 0.0, nrows);
    };
    for (int i = threadIdx.x; i < nrows; i += BLOCK) {
        lv[i] = a.lam_v[row0 + i];
        if (SYM) zs[i] = 0.0;
    }
    __syncthreads();
    for (int k = a.n_steps - 1; k >= 0; --k) {
        const int64_t gstep = p.step0 + k;
        const float* rec = a.record + (int64_t)k * 5 * NR;
        const float* nxt = rec + 5 * NR;  // state at the start of the following step
        if (threadIdx.x == 0) {
            const double vs = nxt[row0 + soma];
            const double r = vs - (double)a.target[(int64_t)cell * a.target_stride + gstep];
            loss += r * r;
            lv[soma] += (float)(2.0 * r);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nrows; i += BLOCK) put_y(i, lv[i], a.inv_c[row0 + i]);
        __syncthreads();
        product();
        __syncthreads();
        // half B and the mechanism
        for (int i = threadIdx.x; i < nrows; i += BLOCK) {
            const int64_t gi = row0 + i;
            const double invc = a.inv_c[gi];
            const double lxB = take_lx(i, invc);
            const double v1 = rec[4 * NR + gi];
            const double mN = nxt[NR + gi], hN = nxt[2 * NR + gi], nN = nxt[3 * NR + gi];
            double lm = a.lam_m[gi], lh = a.lam_h[gi], ln = a.lam_n[gi];
            double gr_na = 0, gr_k = 0, gr_l = 0;
            double lv1 = half_adjoint(lxB, v1, mN, hN, nN, a.g_na[gi], a.g_k[gi], a.g_l[gi], e_na,
                                      e_k, e_l, hs, invc, lm, lh, ln, gr_na, gr_k, gr_l);
            const double g0[3] = {rec[NR + gi], rec[2 * NR + gi], rec[3 * NR + gi]};
            double lg[3] = {lm, lh, ln};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                GateLookup L;
                lookup6(a.table6 + (int64_t)q * p.num_nodes * 6, p.num_nodes, v1, vmin, frq, L);
                lv1 += (L.slope[0] + L.slope[1] * g0[q]) * lg[q];
                th[2 * q] += (L.val[2] + L.val[3] * g0[q]) * lg[q];
                th[2 * q + 1] += (L.val[4] + L.val[5] * g0[q]) * lg[q];
                lg[q] *= L.val[1];
            }
            a.lam_m[gi] = (float)lg[0];
            a.lam_h[gi] = (float)lg[1];
            a.lam_n[gi] = (float)lg[2];
            a.grad_na[gi] += gr_na;
            a.grad_k[gi] += gr_k;
            a.grad_l[gi] += gr_l;
            put_y(i, lv1, invc);
        }
        __syncthreads();
        product();
        __syncthreads();
        // half A with the gates at the start of the step
        for (int i = threadIdx.x; i < nrows; i += BLOCK) {
            const int64_t gi = row0 + i;
            const double invc = a.inv_c[gi];
            const double lxA = take_lx(i, invc);
            double lvk = lxA;
            if (gstep > 0 || p.conductances_valid) {
                double lm = a.lam_m[gi], lh = a.lam_h[gi], ln = a.lam_n[gi];
                double gr_na = 0, gr_k = 0, gr_l = 0;
                lvk = half_adjoint(lxA, rec[gi], rec[NR + gi], rec[2 * NR + gi], rec[3 * NR + gi],
                                   a.g_na[gi], a.g_k[gi], a.g_l[gi], e_na, e_k, e_l, hs, invc, lm,
                                   lh, ln, gr_na, gr_k, gr_l);
                a.lam_m[gi] = (float)lm;
                a.lam_h[gi] = (float)lh;
                a.lam_n[gi] = (float)ln;
                a.grad_na[gi] += gr_na;
                a.grad_k[gi] += gr_k;
                a.grad_l[gi] += gr_l;
            }
            lv[i] = (float)lvk;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nrows; i += BLOCK) a.lam_v[row0 + i] = lv[i];
    // reduce the rate-parameter gradients of the block (fixed order: deterministic)
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double s = wave_sum(th[q]);
        if (lane == 0) red[q][wave] = s;
    }
    if (lane == 0) red_loss[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 6; ++q) {
            double t = 0.0;
            for (int w = 0; w < BLOCK / kWave; ++w) t += red[q][w];
            a.grad_theta[(int64_t)cell * 6 + q] += t;
        }
        double t = 0.0;
        for (int w = 0; w < BLOCK / kWave; ++w) t += red_loss[w];
        a.loss[cell] += t;
    }
}

// ----------------------------------------------------------------------------------------------
// Reverse pass on the Chebyshev store: one launch per half step, in reverse order (HALF 1: the
// loss of step k, P^T through the second half step and the mechanism; HALF 0: P^T through the
// first half step).  P^T lambda = C P (C^-1 lambda) (B = C^-1 G with G symmetric), the same series
// as the forward product.  The voltage adjoint passes between launches as FP64 (lam64); a call
// reads it from lam_v at its first launch and leaves it there after its last.
// ----------------------------------------------------------------------------------------------
template <int HALF, int MAXJ, int NT = kChebBlock, int KIND = kProdAny>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MAXJ == 1 ? 8 : 4)))
k_hh_adj_cheb(AdjArgs a, ChebStore cs, double* lam64, int k, int first, int last) {
    __shared__ __align__(16) double b0[prod_lds(MAXJ, KIND)];
    __shared__ double red[7][NT / kWave];
    const nw_hh_params& p = a.prm;
    const int cell = a.order[blockIdx.x];
    const int64_t row0 = a.cell_row0[cell];
    const int nrows = (int)(a.cell_row0[cell + 1] - row0);
    const int64_t pre_base = cs.pre_base[cell];
    const bool pre = pre_base >= 0;
    const uint16_t* pos = cs.pre_pos + row0;
    const int64_t NR = a.num_rows;
    const int64_t gstep = p.step0 + k;
    const float* rec = a.record + (int64_t)k * 5 * NR;
    const float* nxt = rec + 5 * NR;  // state at the start of the following step
    const int soma = a.soma_local[cell];
    double loss = 0.0;
    for (int i = threadIdx.x; i < nrows; i += NT) {
        const int64_t gi = row0 + i;
        double lam = first ? (double)a.lam_v[gi] : lam64[gi];
        if (HALF == 1 && i == soma) {
            const double r = (double)nxt[gi] - (double)a.target[(int64_t)cell * a.target_stride + gstep];
            loss = r * r;
            lam += 2.0 * r;
        }
        b0[pre ? pos[i] : i] = lam * (double)a.inv_c[gi];
    }
    __syncthreads();
    double out[kChebRows];
    const int places = cheb_cell<MAXJ, NT, KIND>(cs, row0, cell, nrows, pre_base, b0, out);
#pragma unroll
    for (int j = 0; j < kChebRows; ++j) {
        const int r = threadIdx.x + j * kChebBlock;
        if (r < places) b0[r] = out[j];
    }
    __syncthreads();
    const double hs = 0.5 * (double)p.dt_ms * 1e-3;
    const double vmin = p.vmin;
    const double frq = (double)(p.num_nodes - 1) / ((double)p.vmax - (double)p.vmin);
    const double e_na = p.e_na, e_k = p.e_k, e_l = p.e_l;
    double th[6] = {0, 0, 0, 0, 0, 0};
    // the thread's rows unrolled (one in the one-group launches): a loop over them let the
    // compiler hoist the constants of exp() out of it into registers that it then spilled
    // (a 512-thread launch serves cells of at most 512 rows: one row per thread)
    constexpr int ROWS = NT == kChebSmallBlock ? 1 : MAXJ * kChebBlock / NT;
#pragma unroll
    for (int jr = 0; jr < ROWS; ++jr) {
        const int i = threadIdx.x + jr * NT;
        if (i >= nrows) break;
        const int64_t gi = row0 + i;
        const double invc = a.inv_c[gi];
        const double lx = b0[pre ? pos[i] : i] / invc;  // (P^T lambda)_i
        double lam_out = lx;
        if (HALF == 1) {  // half B and the mechanism
            const double v1 = rec[4 * NR + gi];
            const double mN = nxt[NR + gi], hN = nxt[2 * NR + gi], nN = nxt[3 * NR + gi];
            double lm = a.lam_m[gi], lh = a.lam_h[gi], ln = a.lam_n[gi];
            double gr_na = 0, gr_k = 0, gr_l = 0;
            lam_out = half_adjoint(lx, v1, mN, hN, nN, a.g_na[gi], a.g_k[gi], a.g_l[gi], e_na, e_k,
                                   e_l, hs, invc, lm, lh, ln, gr_na, gr_k, gr_l);
            // one gate at a time (fewer live registers), the gate's values rotating through
            // scalars: an array indexed by the loop counter would live in scratch memory
            double ga = rec[NR + gi], gb = rec[2 * NR + gi], gc = rec[3 * NR + gi];
            double la = lm, lb = lh, lc = ln;
            double t0 = th[0], t1 = th[1], t2 = th[2], t3 = th[3], t4 = th[4], t5 = th[5];
#pragma unroll 1
            for (int q = 0; q < 3; ++q) {
                GateLookup L;
                lookup6(a.table6 + (int64_t)q * p.num_nodes * 6, p.num_nodes, v1, vmin, frq, L);
                lam_out += (L.slope[0] + L.slope[1] * ga) * la;
                t0 += (L.val[2] + L.val[3] * ga) * la;
                t1 += (L.val[4] + L.val[5] * ga) * la;
                la *= L.val[1];
                // rotate: the next gate's values move to the front; after three turns every
                // value is back in its place
                const double g_ = ga, l_ = la, u0 = t0, u1 = t1;
                ga = gb; gb = gc; gc = g_;
                la = lb; lb = lc; lc = l_;
                t0 = t2; t1 = t3; t2 = t4; t3 = t5; t4 = u0; t5 = u1;
            }
            th[0] = t0; th[1] = t1; th[2] = t2; th[3] = t3; th[4] = t4; th[5] = t5;
            a.lam_m[gi] = (float)la;
            a.lam_h[gi] = (float)lb;
            a.lam_n[gi] = (float)lc;
            a.grad_na[gi] += gr_na;
            a.grad_k[gi] += gr_k;
            a.grad_l[gi] += gr_l;
        } else if (gstep > 0 || p.conductances_valid) {  // half A with the gates at the start
            double lm = a.lam_m[gi], lh = a.lam_h[gi], ln = a.lam_n[gi];
            double gr_na = 0, gr_k = 0, gr_l = 0;
            lam_out = half_adjoint(lx, rec[gi], rec[NR + gi], rec[2 * NR + gi], rec[3 * NR + gi],
                                   a.g_na[gi], a.g_k[gi], a.g_l[gi], e_na, e_k, e_l, hs, invc, lm,
                                   lh, ln, gr_na, gr_k, gr_l);
            a.lam_m[gi] = (float)lm;
            a.lam_h[gi] = (float)lh;
            a.lam_n[gi] = (float)ln;
            a.grad_na[gi] += gr_na;
            a.grad_k[gi] += gr_k;
            a.grad_l[gi] += gr_l;
        }
        lam64[gi] = lam_out;
        if (last) a.lam_v[gi] = (float)lam_out;
    }
    if (HALF == 1) {  // the block's rate-parameter gradients and loss (fixed order: deterministic)
        const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double t = wave_sum(th[q]);
            if (lane == 0) red[q][wave] = t;
        }
        const double l = wave_sum(loss);
        if (lane == 0) red[6][wave] = l;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 0; q < 7; ++q) {
                double t = 0.0;
                for (int w = 0; w < NT / kWave; ++w) t += red[q][w];
                if (q < 6) a.grad_theta[(int64_t)cell * 6 + q] += t;
                else a.loss[cell] += t;
            }
        }
    }
}

// ----------------------------------------------------------------------------------------------
// HH mechanism instances (generic staggered path): thread per instance, voltage gathered through
// the instance's segment; writes the three conductances interleaved (na, k, l).
// ----------------------------------------------------------------------------------------------
template <int GATE_MODE>
__global__ void __launch_bounds__(256)
k_hh_mechanism(int64_t M, double dt, double q10, const float* __restrict__ table, int num_nodes,
               double vmin, double vmax, const float* __restrict__ voltage,
               const int32_t* __restrict__ seg_index, float* m_, float* h_, float* n_,
               const float* __restrict__ gna, const float* __restrict__ gk,
               const float* __restrict__ gl, float* __restrict__ g_out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double v = voltage[seg_index ? (int64_t)seg_index[i] : i];
    double m = m_[i], h = h_[i], n = n_[i];
    if (GATE_MODE == 1) {
        hh_exact(v, q10, dt, m, h, n);
    } else {
        const int nb = num_nodes - 1;
        const double frq = (double)nb / (vmax - vmin);
        int b;
        double f;
        bucket_of(v, vmin, frq, nb, b, f);
        const float2* tm = reinterpret_cast<const float2*>(table);
        m = gate_advance(tm, b, f, m);
        h = gate_advance(tm + num_nodes, b, f, h);
        n = gate_advance(tm + 2 * num_nodes, b, f, n);
    }
    const float m32 = (float)m, h32 = (float)h, n32 = (float)n;
    m_[i] = m32;
    h_[i] = h32;
    n_[i] = n32;
    const double mm = m32, hh = h32, nn = n32;
    g_out[3 * i + 0] = (float)((double)gna[i] * mm * mm * mm * hh);
    g_out[3 * i + 1] = (float)((double)gk[i] * nn * nn * nn * nn);
    g_out[3 * i + 2] = gl[i];
}

// ----------------------------------------------------------------------------------------------
// LTI kinetic kernel: thread per instance, S states in registers.
// ----------------------------------------------------------------------------------------------
template <int S, bool HOLD>
__global__ void __launch_bounds__(256)
k_lti(int num_nodes, float vmin, float vmax, const float* __restrict__ table, int64_t M,
      const float* __restrict__ voltage, const int32_t* __restrict__ seg_index, float* states,
      int64_t stride, float conserve_sum, int n_steps, const float* __restrict__ weights,
      const float* __restrict__ gmax, float* __restrict__ cond_out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int nb = num_nodes - 1;
    const double bucket_frq = (double)nb / ((double)vmax - (double)vmin);
    const double v = voltage[seg_index ? (int64_t)seg_index[i] : i];
    int b;
    double fd;
    bucket_of<double>(v, (double)vmin, bucket_frq, nb, b, fd);
    const float* T0 = table + (int64_t)b * S * S;
    const float* T1 = T0 + S * S;
    // States and sums in FP64 registers, propagator entries in FP32 (as stored).
    double s[S];
#pragma unroll
    for (int k = 0; k < S; ++k) s[k] = states[k * stride + i];
    // With the input held constant over the launch (HOLD) the interpolated propagator is formed
    // once and kept in registers; every step is then identical to a call with n_steps = 1.
    float Pm[HOLD ? S * S : 1];
    if (HOLD) {
#pragma unroll
        for (int q = 0; q < S * S; ++q)
            Pm[q] = (float)fma(fd, (double)T1[q] - (double)T0[q], (double)T0[q]);
    }
    for (int step = 0; step < n_steps; ++step) {
        double out[S];
#pragma unroll
        for (int r = 0; r < S; ++r) out[r] = 0.0;
#pragma unroll
        for (int c = 0; c < S; ++c) {  // column-major walk, as the reference kernel (codegen.py:81-94)
#pragma unroll
            for (int r = 0; r < S; ++r) {
                double pe;
                if (HOLD) {
                    pe = Pm[r * S + c];
                } else {
                    const double t0 = T0[r * S + c];
                    pe = fma(fd, (double)T1[r * S + c] - t0, t0);
                }
                out[r] = fma(pe, s[c], out[r]);
            }
        }
        if (conserve_sum > 0.0f) {  // codegen.py:95-102
            double tot = 0.0;
#pragma unroll
            for (int r = 0; r < S; ++r) tot += out[r];
            const double corr = (double)conserve_sum / tot;
#pragma unroll
            for (int r = 0; r < S; ++r) out[r] *= corr;
        }
#pragma unroll
        for (int r = 0; r < S; ++r) s[r] = out[r];
    }
#pragma unroll
    for (int k = 0; k < S; ++k) states[k * stride + i] = (float)s[k];
    if (cond_out) {
        double w = 0.0;
#pragma unroll
        for (int k = 0; k < S; ++k) w = fma((double)weights[k], (double)(float)s[k], w);
        cond_out[i] = (float)((double)gmax[i] * w);
    }
}

// ----------------------------------------------------------------------------------------------
// LTI kinetic advance with K = 1 or 2 inputs (linear or log2 mapping), thread per instance.
// ----------------------------------------------------------------------------------------------
struct LtiInputs {
    nw_lti_input in[2];
};

__device__ __forceinline__ void input_node(const nw_lti_input& q, int64_t i, int& b, double& f) {
    const double x = (double)q.values[q.index ? (int64_t)q.index[i] : i];
    const int nb = q.num_nodes - 1;
    double u;
    if (q.mapping == NW_INPUT_LOG2) {
        const double lo = log2(q.minimum + q.scale), hi = log2(q.maximum + q.scale);
        u = (log2(x + q.scale) - lo) * ((double)nb / (hi - lo));
    } else {
        u = (x - q.minimum) * ((double)nb / (q.maximum - q.minimum));
    }
    // int() of the reference truncates; NaN (a negative concentration below -scale) and values
    // below the range take node 0 with fraction 0
    if (!(u >= 0.0)) {
        b = 0;
        f = 0.0;
    } else if (u >= (double)nb) {
        b = nb - 1;
        f = 1.0;
    } else {
        b = (int)u;
        f = u - (double)b;
    }
}

template <int S, int K>
__global__ __launch_bounds__(256) void
k_lti_nd(LtiInputs q, const float* __restrict__ table, int64_t M, float* states, int64_t stride,
         float conserve_sum, int n_steps, const float* __restrict__ weights,
         const float* __restrict__ gmax, float* __restrict__ cond_out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= M) return;
    int b0, b1 = 0;
    double f0, f1 = 0.0;
    input_node(q.in[0], i, b0, f0);
    if (K == 2) input_node(q.in[1], i, b1, f1);
    const int n1 = K == 2 ? q.in[1].num_nodes : 1;
    // corner weights of the (bi)linear interpolation
    const float* T00 = table + ((int64_t)b0 * n1 + b1) * S * S;
    const float* T10 = T00 + (int64_t)n1 * S * S;      // next node of input 0
    const float* T01 = T00 + S * S;                     // next node of input 1
    const float* T11 = T10 + S * S;
    const double w00 = (1.0 - f0) * (K == 2 ? 1.0 - f1 : 1.0), w10 = f0 * (K == 2 ? 1.0 - f1 : 1.0);
    const double w01 = K == 2 ? (1.0 - f0) * f1 : 0.0, w11 = K == 2 ? f0 * f1 : 0.0;
    double s[S];
#pragma unroll
    for (int k = 0; k < S; ++k) s[k] = states[k * stride + i];
    for (int step = 0; step < n_steps; ++step) {
        double out[S];
#pragma unroll
        for (int r = 0; r < S; ++r) out[r] = 0.0;
#pragma unroll
        for (int c = 0; c < S; ++c) {
#pragma unroll
            for (int r = 0; r < S; ++r) {
                const int e = r * S + c;
                double pe = w00 * (double)T00[e] + w10 * (double)T10[e];
                if (K == 2) pe += w01 * (double)T01[e] + w11 * (double)T11[e];
                out[r] = fma(pe, s[c], out[r]);
            }
        }
        if (conserve_sum > 0.0f) {  // codegen.py:95-102
            double tot = 0.0;
#pragma unroll
            for (int r = 0; r < S; ++r) tot += out[r];
            const double corr = (double)conserve_sum / tot;
#pragma unroll
            for (int r = 0; r < S; ++r) out[r] *= corr;
        }
#pragma unroll
        for (int r = 0; r < S; ++r) s[r] = out[r];
    }
#pragma unroll
    for (int k = 0; k < S; ++k) states[k * stride + i] = (float)s[k];
    if (cond_out) {
        double w = 0.0;
#pragma unroll
        for (int k = 0; k < S; ++k) w = fma((double)weights[k], (double)(float)s[k], w);
        cond_out[i] = (float)((double)gmax[i] * w);
    }
}

// ----------------------------------------------------------------------------------------------
// Segmented accumulation: thread per segment over its instances (sorted by segment).  The
// instances are taken four at a time: their indices and reversal potentials load together, then
// the four gathered conductances, so a segment with up to four sources costs two memory round
// trips instead of two per source; the sums run in instance order, as a serial loop's.  Where the
// four slots lie inside the arrays their reversal potentials and indices are 16-byte loads (the
// slots past the segment's last belong to the next segment and are not summed), and so are the
// four conductances where their indices run consecutively (the model's interleaved mechanism
// output): 5.66 against 5.10 TB/s with one load per slot.  (Staging a block's sources in LDS with consecutive threads
// on consecutive sources measured slower: 4.35 against 5.10 TB/s on 10M segments of three
// sources.)
// ----------------------------------------------------------------------------------------------
constexpr int kAccBlock = 256;

__global__ void __launch_bounds__(kAccBlock)
k_accumulate(int64_t num_segments, const int64_t* __restrict__ ptr,
             const float* __restrict__ g_inst, const int64_t* __restrict__ g_index,
             const float* __restrict__ e_inst, const float* __restrict__ nonspec_i,
             float* __restrict__ sum_g, float* __restrict__ drive_v, float* __restrict__ sum_i) {
    const int64_t s = (int64_t)blockIdx.x * kAccBlock + threadIdx.x;
    if (s >= num_segments) return;
    double g = 0.0, num = 0.0;
    const int64_t k1 = ptr[s + 1];
    const int64_t total = ptr[num_segments];
    for (int64_t k0 = ptr[s]; k0 < k1; k0 += 4) {
        int64_t ix[4];
        float ev[4], gv[4];
        if (k0 + 4 <= total) {
            // four adjacent slots read unconditionally (wide loads; the slots past k1 belong to
            // the next segment and are not summed)
            typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
            typedef long long l4u __attribute__((ext_vector_type(4), aligned(8)));
            const f4u e4 = *reinterpret_cast<const f4u*>(e_inst + k0);
            ev[0] = e4.x; ev[1] = e4.y; ev[2] = e4.z; ev[3] = e4.w;
            if (g_index) {
                const l4u i4 = *reinterpret_cast<const l4u*>(g_index + k0);
                ix[0] = i4.x; ix[1] = i4.y; ix[2] = i4.z; ix[3] = i4.w;
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) ix[q] = k0 + q;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t k = k0 + q;
                ix[q] = 0;
                ev[q] = 0.0f;
                if (k < k1) {
                    ix[q] = g_index ? g_index[k] : k;
                    ev[q] = e_inst[k];
                }
            }
        }
        const bool run = k0 + 4 <= total && ix[1] == ix[0] + 1 && ix[2] == ix[0] + 2 &&
                         ix[3] == ix[0] + 3;
        if (run) {
            // four consecutive conductances (the last one's index is read from the index array,
            // so all four lie in bounds): one 16-byte load
            typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
            const f4u g4 = *reinterpret_cast<const f4u*>(g_inst + ix[0]);
            gv[0] = g4.x; gv[1] = k0 + 1 < k1 ? g4.y : 0.0f;
            gv[2] = k0 + 2 < k1 ? g4.z : 0.0f; gv[3] = k0 + 3 < k1 ? g4.w : 0.0f;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) gv[q] = (k0 + q < k1) ? g_inst[ix[q]] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (k0 + q < k1) {
                const double gi = gv[q];
                g += gi;
                num = fma(gi, (double)ev[q], num);
            }
        }
    }
    sum_g[s] = (float)g;
    drive_v[s] = (g != 0.0) ? (float)(num / g) : 0.0f;
    sum_i[s] = nonspec_i ? nonspec_i[s] : 0.0f;
}

// ----------------------------------------------------------------------------------------------
// Intracellular Ca2+ with a first-order buffer and diffusion (species.py:54-70 with a diffusion
// propagator; buffer B <-> CaB, forward kon*Ca, backward koff, CONSERVE B + CaB = btot).
// Per step: c = P_D max(0, c + dcdt dt/2); buffer over dt at c(t+dt/2), dcdt = -(CaB' - CaB)/dt;
// c = P_D max(0, c + dcdt dt/2).  One workgroup per cell, concentrations stay in LDS.  The bound
// buffer is stored as an FP32 offset from the scalar cab_ref: it is large next to free Ca, and its
// per-step change must keep the precision of the free concentration.
// ----------------------------------------------------------------------------------------------
struct CaArgs {
    const int64_t* cell_row0;
    const int64_t* cell_nz0;
    const uint32_t* row_off;
    const float* vals;
    const uint16_t* cols;
    const double* rowsum;
    const int32_t* order;
    int32_t nmax;
    int32_t n_steps;
    nw_ca_params prm;
    float* ca;
    float* cab;
    float* dcdt;
};

template <int BLOCK, int ROWMODE>
__global__ void __launch_bounds__(BLOCK) k_ca_step(CaArgs a) {
    extern __shared__ float lds[];
    float* xs = lds;
    float* cs = lds + a.nmax;
    const int cell = a.order[blockIdx.x];
    const int64_t row0 = a.cell_row0[cell];
    const int nrows = (int)(a.cell_row0[cell + 1] - row0);
    const uint32_t* roff = a.row_off + row0 + cell;
    const float* vals = a.vals + a.cell_nz0[cell];
    const uint16_t* cols = a.cols + a.cell_nz0[cell];
    const double* rs = a.rowsum + row0;
    const double dt = a.prm.dt_ms, hdt = 0.5 * dt;
    const double kon = a.prm.kon, koff = a.prm.koff, btot = a.prm.btot;
    const double cab_ref = a.prm.cab_ref;
    for (int i = threadIdx.x; i < nrows; i += BLOCK) cs[i] = a.ca[row0 + i];
    __syncthreads();
    for (int step = 0; step < a.n_steps; ++step) {
        for (int i = threadIdx.x; i < nrows; i += BLOCK)
            xs[i] = (float)fmax(0.0, (double)cs[i] + (double)a.dcdt[row0 + i] * hdt);
        __syncthreads();
        spmv<ROWMODE>(roff, vals, cols, rs, xs, cs, 0.0, nrows);
        __syncthreads();
        for (int i = threadIdx.x; i < nrows; i += BLOCK) {
            const int64_t gi = row0 + i;
            const double c = cs[i];
            const float stored = a.cab[gi];
            const double cab = cab_ref + (double)stored;
            const double fw = kon * c, s = fw + koff;
            const double inf = btot * fw / s;
            const float stored_new = (float)(inf + (cab - inf) * exp(-s * dt) - cab_ref);
            // the Ca change is the change of the stored bound amount: mass stays exact
            const double d = -((double)stored_new - (double)stored) / dt;
            a.cab[gi] = stored_new;
            a.dcdt[gi] = (float)d;
            xs[i] = (float)fmax(0.0, c + d * hdt);
        }
        __syncthreads();
        spmv<ROWMODE>(roff, vals, cols, rs, xs, cs, 0.0, nrows);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nrows; i += BLOCK) a.ca[row0 + i] = cs[i];
}

// out[b] = sum over this block's slice of w * (x + y + z_scale * z), FP64 (mass checks).
__global__ void k_weighted_sum(int64_t n, const double* __restrict__ w, const float* __restrict__ x,
                               const float* __restrict__ y, const float* __restrict__ z,
                               double z_scale, double* __restrict__ out) {
    __shared__ double part[256 / kWave];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        acc += w[i] * ((double)x[i] + (y ? (double)y[i] : 0.0) + (z ? z_scale * (double)z[i] : 0.0));
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x / kWave); ++k) t += part[k];
        out[blockIdx.x] = t;
    }
}

// total[0] = part[0] + part[1] + ... + part[num - 1], summed in that order by one thread (the
// order of a host loop over the partials, so the two agree bit for bit).
__global__ void k_sum_in_order(const double* __restrict__ part, int num, double* __restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < num; ++k) t += part[k];
        total[0] = t;
    }
}

// AP detector of every neuron (neuron.py:72-76): over = v[seg] >= threshold;
// detected = over and not previously over; state = over.
__global__ void k_ap_detect(int64_t num, const int32_t* __restrict__ seg,
                            const float* __restrict__ v, float threshold, uint8_t* state,
                            uint8_t* detected) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= num) return;
    const bool over = seg[i] >= 0 && v[seg[i]] >= threshold;
    detected[i] = (over && !state[i]) ? 1 : 0;
    state[i] = over ? 1 : 0;
}

__global__ void k_set_i64(int64_t* __restrict__ dst, int64_t value) {
    if (threadIdx.x == 0) *dst = value;
}

// The caller-provided arguments of a call (every one but prm.step0) as bytes: the key of a
// graph.  Arguments derived from the handle are the same for every call on it.
static std::vector<unsigned char> hh_step_key(const nw_hh_params& prm, int n_steps, const HHArgs& a) {
    static_assert(sizeof(nw_hh_params) == 88, "nw_hh_params has no padding");
    nw_hh_params q = prm;
    q.step0 = 0;
    const void* ptrs[] = {a.table, a.v, a.m, a.h, a.n, a.g_na, a.g_k, a.g_l, a.inv_c, a.stim_local,
                          a.ap_local, a.ap_state, a.ap_count, a.trace, a.record};
    std::vector<unsigned char> key(sizeof q + sizeof ptrs + 2 * sizeof(int64_t));
    unsigned char* k = key.data();
    std::memcpy(k, &q, sizeof q);
    std::memcpy(k + sizeof q, ptrs, sizeof ptrs);
    const int64_t tail[2] = {n_steps, a.trace_stride};
    std::memcpy(k + sizeof q + sizeof ptrs, tail, sizeof tail);
    return key;
}

// NW_LAUNCH_GRAPH / NW_LAUNCH_GRAPH_PREPARE: find or capture the graph of this argument set,
// then (GRAPH) write the first step's index to the handle's device word and replay the graph.
template <typename Issue>
static int hh_step_graph(const nw_propagator* p, const nw_hh_params& prm, int n_steps, HHArgs a,
                         Issue& issue, hipStream_t s) {
    NW_REQUIRE(p->step_dev, "the store has no device step counter");
    a.step_dev = p->step_dev;
    std::vector<unsigned char> key = hh_step_key(prm, n_steps, a);
    nw_propagator::StepGraph* hit = nullptr;
    for (auto& g : p->graphs)
        if (g.key == key) hit = &g;
    if (!hit) {
        if (!p->capture_stream)
            NW_HIP(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        NW_HIP(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
        issue(a, p->capture_stream);
        const hipError_t launch_err = hipGetLastError();
        const hipError_t end_err = hipStreamEndCapture(p->capture_stream, &graph);
        NW_HIP(launch_err);
        NW_HIP(end_err);
        hipGraphExec_t exec = nullptr;
        const hipError_t inst_err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        NW_HIP(inst_err);
        if ((int)p->graphs.size() >= kStepGraphs) {  // drop the least recently used
            size_t old = 0;
            for (size_t i = 1; i < p->graphs.size(); ++i)
                if (p->graphs[i].used < p->graphs[old].used) old = i;
            (void)hipGraphExecDestroy(p->graphs[old].exec);
            p->graphs.erase(p->graphs.begin() + (std::ptrdiff_t)old);
        }
        p->graphs.push_back(nw_propagator::StepGraph{std::move(key), exec, 0});
        hit = &p->graphs.back();
    }
    hit->used = ++p->graph_clock;
    if (prm.launch_mode == NW_LAUNCH_GRAPH_PREPARE) return NW_OK;
    k_set_i64<<<1, 64, 0, s>>>(p->step_dev, prm.step0);
    NW_HIP(hipGetLastError());
    NW_HIP(hipGraphLaunch(hit->exec, s));
    return NW_OK;
}

}  // namespace nw

// Partials of nw_weighted_sum_total: a file-scope device array, one copy per device, held until
// the process ends.  Host code reaches it through the address hipGetSymbolAddress returns for the
// current device (scratch_of_device), never through the array's name.
constexpr int kSumScratch = NW_SUM_SCRATCH_PARTIALS;
static __device__ __align__(256) double scratch[kSumScratch];

namespace nw {

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the first call there.
static int scratch_of_device(double** out) {
    int dev = 0;
    NW_HIP(hipGetDevice(&dev));
    NW_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev,
               kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        NW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NW_REQUIRE(p, "no device address for the partials array on device %d", dev);
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return NW_OK;
}

}  // namespace nw

using namespace nw;

extern "C" {

int nw_hh_step(const nw_propagator* p, const nw_hh_params* prm, int32_t n_steps,
               const float* gate_table, float* v, float* m, float* h, float* n, const float* g_na,
               const float* g_k, const float* g_l, const float* inv_c, const int32_t* stim_local,
               const int32_t* ap_local, uint8_t* ap_state, uint32_t* ap_count, float* trace,
               int64_t trace_stride, float* record, void* stream) {
    NW_REQUIRE(p && prm && gate_table && v && m && h && n && g_na && g_k && g_l && inv_c,
               "nw_hh_step: null argument");
    NW_REQUIRE(n_steps >= 0, "n_steps must be non-negative");
    NW_REQUIRE(prm->num_nodes >= 2 && prm->vmax > prm->vmin, "bad gate table grid");
    NW_REQUIRE(!ap_local || (ap_state && ap_count), "AP detector needs ap_state and ap_count");
    NW_REQUIRE(prm->stim_mode == 0 || stim_local, "stimulus needs stim_local");
    NW_REQUIRE(prm->stim_mode != 2 || (prm->pulse_ticks >= 0 && prm->pulse_ticks <= 100000),
               "pulse_ticks out of range");
    if (n_steps == 0) return NW_OK;
    HHArgs a{p->cell_row0, p->cell_nz0, p->row_off, p->vals, p->cols, p->rowsum, p->order,
             p->max_cell_rows,
             n_steps, *prm, gate_table, v, m, h, n, g_na, g_k, g_l, inv_c, stim_local, ap_local,
             ap_state, ap_count, trace, trace_stride, record, p->num_rows, p->wgt};
    a.cheb = cheb_store(p);
    const bool sym = p->symmetric != 0;
    const size_t lds = (p->chebyshev ? 2 * sizeof(double)
                                     : sym ? sizeof(double) + sizeof(float) : 2 * sizeof(float)) *
                       (size_t)p->max_cell_rows;
    NW_REQUIRE(lds <= 160 * 1024, "cell of %d segments exceeds LDS", p->max_cell_rows);
    hipStream_t s = (hipStream_t)stream;
    NW_REQUIRE(prm->gate_mode == 0 || prm->gate_mode == 1, "gate_mode must be 0 or 1");
    if (p->chebyshev) {
        NW_REQUIRE(prm->block_threads == 0 || prm->block_threads == kChebBlock,
                   "the Chebyshev store runs %d threads per workgroup", kChebBlock);
        NW_REQUIRE(p->max_cell_rows <= kChebBuf, "cell of %d segments exceeds the Chebyshev store",
                   p->max_cell_rows);
        const unsigned grid = (unsigned)p->num_cells;
        double* v64 = p->cheb_v64;
        NW_REQUIRE(prm->launch_mode >= NW_LAUNCH_DIRECT && prm->launch_mode <= NW_LAUNCH_GRAPH_PREPARE,
                   "launch_mode must be NW_LAUNCH_DIRECT, NW_LAUNCH_GRAPH or NW_LAUNCH_GRAPH_PREPARE");
        NW_REQUIRE(prm->cheb_launch >= NW_CHEB_LAUNCH_DEFAULT &&
                       prm->cheb_launch <= NW_CHEB_LAUNCH_TWO_STEPS,
                   "cheb_launch must be one of NW_CHEB_LAUNCH_*");
        // one launch per half step, built for the store's product and the largest cell
        auto issue = [&](const HHArgs& args, hipStream_t st_) {
            auto launch = [&](auto gate, auto maxj, auto nt, auto kind) {
                constexpr int G = decltype(gate)::value, MJ = decltype(maxj)::value;
                constexpr int T = decltype(nt)::value, KD = decltype(kind)::value;
                // the Chebyshev store by default: two steps per launch (the last step of an odd
                // count alone) where a workgroup holds the CU (five row groups), one step per
                // launch where two to four share it (one row group); the rational store, whose
                // fused build spills, one launch per half step
                int form = NW_CHEB_LAUNCH_HALF;
                if constexpr (KD == kProdChebyshev) {
                    form = args.prm.cheb_launch != NW_CHEB_LAUNCH_DEFAULT ? args.prm.cheb_launch
                           : MJ > 1 ? NW_CHEB_LAUNCH_TWO_STEPS : NW_CHEB_LAUNCH_STEP;
                }
                int st = 0;
                if constexpr (KD == kProdChebyshev) {
                    if (form == NW_CHEB_LAUNCH_TWO_STEPS)
                        for (; st + 2 <= n_steps; st += 2)
                            k_hh_step2_cheb<G, MJ, T, KD><<<grid, T, 0, st_>>>(args, v64, st, st == 0);
                    if (form != NW_CHEB_LAUNCH_HALF)
                        for (; st < n_steps; ++st)
                            k_hh_step_cheb<G, MJ, T, KD><<<grid, T, 0, st_>>>(args, v64, st, st == 0);
                }
                for (; st < n_steps; ++st) {
                    k_hh_half_cheb<G, 0, MJ, T, KD><<<grid, T, 0, st_>>>(args, v64, st, st == 0);
                    k_hh_half_cheb<G, 1, MJ, T, KD><<<grid, T, 0, st_>>>(args, v64, st, 0);
                }
            };
            auto by_shape = [&](auto gate, auto kind) {
                using I = std::integral_constant<int, 1>;
                using I5 = std::integral_constant<int, kChebRows>;
                using TS = std::integral_constant<int, kChebSmallBlock>;
                using TB = std::integral_constant<int, kChebBlock>;
                if (p->cheb_small) launch(gate, I{}, TS{}, kind);
                else if (p->cheb_maxj == 1) launch(gate, I{}, TB{}, kind);
                else launch(gate, I5{}, TB{}, kind);
            };
            auto by_kind = [&](auto gate) {
                if (p->rational) by_shape(gate, std::integral_constant<int, kProdRational>{});
                else by_shape(gate, std::integral_constant<int, kProdChebyshev>{});
            };
            if (prm->gate_mode == 1) by_kind(std::integral_constant<int, 1>{});
            else by_kind(std::integral_constant<int, 0>{});
        };
        bool graph = prm->launch_mode != NW_LAUNCH_DIRECT;
        if (graph) {
            hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
            NW_HIP(hipStreamIsCapturing(s, &cap));
            if (cap != hipStreamCaptureStatusNone) {
                NW_REQUIRE(prm->launch_mode != NW_LAUNCH_GRAPH_PREPARE,
                           "NW_LAUNCH_GRAPH_PREPARE on a stream that is being captured");
                graph = false;
            }
        }
        if (!graph) {
            issue(a, s);
            NW_HIP(hipGetLastError());
            return NW_OK;
        }
        return hh_step_graph(p, *prm, n_steps, a, issue, s);
    }
    // default: 512 threads for the symmetric format (two workgroups per CU for 5000-segment cells
    // under the LDS budget), 256 for the general one (flat from 256 to 1024 at 10M compartments)
    const int bt = prm->block_threads ? prm->block_threads : (sym ? 512 : 256);
    const unsigned grid = (unsigned)p->num_cells;
#define NW_HH_LAUNCH(B)                                                                   \
    case B:                                                                               \
        if (sym) {                                                                        \
            if (prm->gate_mode == 1) k_hh_step<B, 1, true><<<grid, B, lds, s>>>(a);       \
            else k_hh_step<B, 0, true><<<grid, B, lds, s>>>(a);                           \
        } else {                                                                          \
            if (prm->gate_mode == 1) k_hh_step<B, 1, false><<<grid, B, lds, s>>>(a);      \
            else k_hh_step<B, 0, false><<<grid, B, lds, s>>>(a);                          \
        }                                                                                 \
        break;
    switch (bt) {
        NW_HH_LAUNCH(128) NW_HH_LAUNCH(256) NW_HH_LAUNCH(512) NW_HH_LAUNCH(1024)
        default: NW_REQUIRE(false, "block_threads must be 128, 256, 512 or 1024 (got %d)", bt);
    }
#undef NW_HH_LAUNCH
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_hh_adjoint(const nw_propagator* p, const nw_hh_params* prm, int32_t n_steps,
                  const float* record, const float* table6, const float* g_na, const float* g_k,
                  const float* g_l, const float* inv_c, const int32_t* soma_local,
                  const float* target, int64_t target_stride, float* lam_v, float* lam_m,
                  float* lam_h, float* lam_n, double* grad_na, double* grad_k, double* grad_l,
                  double* grad_theta, double* loss, void* stream) {
    NW_REQUIRE(p && prm && record && table6 && g_na && g_k && g_l && inv_c && soma_local &&
                   target && lam_v && lam_m && lam_h && lam_n && grad_na && grad_k && grad_l &&
                   grad_theta && loss,
               "nw_hh_adjoint: null argument");
    NW_REQUIRE(prm->gate_mode == 0, "the adjoint is written for the table gate mode");
    NW_REQUIRE(prm->stim_mode == 0 || prm->stim_mode == 1, "adjoint supports constant stimulus only");
    NW_REQUIRE(n_steps >= 0 && prm->num_nodes >= 2, "bad arguments");
    if (n_steps == 0) return NW_OK;
    AdjArgs a{p->cell_row0, p->cell_nz0, p->row_off, p->vals, p->cols, p->rowsum, p->order,
              p->max_cell_rows, n_steps, *prm, record, p->num_rows, table6, g_na, g_k, g_l, inv_c,
              soma_local, target, target_stride, lam_v, lam_m, lam_h, lam_n, grad_na, grad_k,
              grad_l, grad_theta, loss, p->wgt};
    if (p->chebyshev) {
        NW_REQUIRE(p->max_cell_rows <= kChebBuf, "cell of %d segments exceeds the Chebyshev store",
                   p->max_cell_rows);
        const ChebStore cs = cheb_store(p);
        const unsigned grid = (unsigned)p->num_cells;
        hipStream_t s = (hipStream_t)stream;
        // one launch per half step in reverse order, built for the store's product and the
        // largest cell
        auto launch = [&](auto maxj, auto nt, auto kind) {
            constexpr int MJ = decltype(maxj)::value, T = decltype(nt)::value;
            constexpr int KD = decltype(kind)::value;
            for (int k = n_steps - 1; k >= 0; --k) {
                k_hh_adj_cheb<1, MJ, T, KD><<<grid, T, 0, s>>>(a, cs, p->cheb_v64, k,
                                                               k == n_steps - 1, 0);
                k_hh_adj_cheb<0, MJ, T, KD><<<grid, T, 0, s>>>(a, cs, p->cheb_v64, k, 0, k == 0);
            }
        };
        auto by_shape = [&](auto kind) {
            using I = std::integral_constant<int, 1>;
            using I5 = std::integral_constant<int, kChebRows>;
            using TS = std::integral_constant<int, kChebSmallBlock>;
            using TB = std::integral_constant<int, kChebBlock>;
            if (p->cheb_small) launch(I{}, TS{}, kind);
            else if (p->cheb_maxj == 1) launch(I{}, TB{}, kind);
            else launch(I5{}, TB{}, kind);
        };
        if (p->rational) by_shape(std::integral_constant<int, kProdRational>{});
        else by_shape(std::integral_constant<int, kProdChebyshev>{});
        NW_HIP(hipGetLastError());
        return NW_OK;
    }
    const bool sym = p->symmetric != 0;
    const size_t lds = (sym ? sizeof(double) + 2 * sizeof(float) : 3 * sizeof(float)) *
                       (size_t)p->max_cell_rows;
    NW_REQUIRE(lds <= 150 * 1024, "cell of %d segments exceeds LDS", p->max_cell_rows);
    if (sym)
        k_hh_adjoint<256, true><<<(unsigned)p->num_cells, 256, lds, (hipStream_t)stream>>>(a);
    else
        k_hh_adjoint<256, false><<<(unsigned)p->num_cells, 256, lds, (hipStream_t)stream>>>(a);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_hh_mechanism(int64_t M, int32_t gate_mode, float dt_ms, float q10, const float* gate_table,
                    int32_t num_nodes, float vmin, float vmax, const float* voltage,
                    const int32_t* seg_index, float* m, float* h, float* n, const float* g_na,
                    const float* g_k, const float* g_l, float* g_out, void* stream) {
    NW_REQUIRE(voltage && m && h && n && g_na && g_k && g_l && g_out, "nw_hh_mechanism: null argument");
    NW_REQUIRE(gate_mode == 1 || (gate_table && num_nodes >= 2 && vmax > vmin), "bad gate table");
    NW_REQUIRE(gate_mode == 0 || gate_mode == 1, "gate_mode must be 0 or 1");
    if (M <= 0) return NW_OK;
    const unsigned grid = (unsigned)((M + 255) / 256);
    hipStream_t s = (hipStream_t)stream;
    if (gate_mode == 1)
        k_hh_mechanism<1><<<grid, 256, 0, s>>>(M, dt_ms, q10, gate_table, num_nodes, vmin, vmax,
                                               voltage, seg_index, m, h, n, g_na, g_k, g_l, g_out);
    else
        k_hh_mechanism<0><<<grid, 256, 0, s>>>(M, dt_ms, q10, gate_table, num_nodes, vmin, vmax,
                                               voltage, seg_index, m, h, n, g_na, g_k, g_l, g_out);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_lti_advance(int32_t S, int32_t num_nodes, float vmin, float vmax, const float* node_table,
                   int64_t M, const float* voltage, const int32_t* seg_index, float* states,
                   int64_t state_stride, float conserve_sum, int32_t n_steps, const float* weights,
                   const float* gmax, float* cond_out, void* stream) {
    NW_REQUIRE(node_table && voltage && states, "nw_lti_advance: null argument");
    NW_REQUIRE(num_nodes >= 2 && vmax > vmin, "bad table grid");
    NW_REQUIRE(M >= 0 && state_stride >= M && n_steps >= 0, "bad sizes");
    NW_REQUIRE(!cond_out || (weights && gmax), "conductance output needs weights and gmax");
    if (M == 0) return NW_OK;
    const unsigned grid = (unsigned)((M + 255) / 256);
    hipStream_t s = (hipStream_t)stream;
#define NW_LTI_CASE(K)                                                                             \
    case K:                                                                                        \
        if (n_steps > 1)                                                                           \
            k_lti<K, true><<<grid, 256, 0, s>>>(num_nodes, vmin, vmax, node_table, M, voltage,     \
                                                seg_index, states, state_stride, conserve_sum,     \
                                                n_steps, weights, gmax, cond_out);                 \
        else                                                                                       \
            k_lti<K, false><<<grid, 256, 0, s>>>(num_nodes, vmin, vmax, node_table, M, voltage,    \
                                                 seg_index, states, state_stride, conserve_sum,    \
                                                 n_steps, weights, gmax, cond_out);                \
        break;
    switch (S) {
        NW_LTI_CASE(1) NW_LTI_CASE(2) NW_LTI_CASE(3) NW_LTI_CASE(4) NW_LTI_CASE(5)
        NW_LTI_CASE(6) NW_LTI_CASE(7) NW_LTI_CASE(8) NW_LTI_CASE(9) NW_LTI_CASE(10)
        NW_LTI_CASE(11) NW_LTI_CASE(12) NW_LTI_CASE(13) NW_LTI_CASE(14) NW_LTI_CASE(15)
        NW_LTI_CASE(16)
        default: NW_REQUIRE(false, "state count %d not supported (1..16)", S);
    }
#undef NW_LTI_CASE
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_lti_advance_nd(int32_t S, int32_t num_inputs, const nw_lti_input* inputs_host,
                      const float* node_table, int64_t M, float* states, int64_t state_stride,
                      float conserve_sum, int32_t n_steps, const float* weights, const float* gmax,
                      float* cond_out, void* stream) {
    NW_REQUIRE(inputs_host && node_table && states, "nw_lti_advance_nd: null argument");
    NW_REQUIRE(num_inputs == 1 || num_inputs == 2, "1 or 2 inputs (got %d)", num_inputs);
    NW_REQUIRE(M >= 0 && state_stride >= M && n_steps >= 0, "bad sizes");
    NW_REQUIRE(!cond_out || (weights && gmax), "conductance output needs weights and gmax");
    LtiInputs q{};
    for (int k = 0; k < num_inputs; ++k) {
        const nw_lti_input& in = inputs_host[k];
        NW_REQUIRE(in.values, "input %d: null values", k);
        NW_REQUIRE(in.num_nodes >= 2 && in.maximum > in.minimum, "input %d: bad grid", k);
        NW_REQUIRE(in.mapping == NW_INPUT_LINEAR ||
                       (in.mapping == NW_INPUT_LOG2 && in.scale > 0.0 && in.minimum + in.scale > 0.0),
                   "input %d: bad mapping", k);
        q.in[k] = in;
    }
    if (M == 0 || n_steps == 0) return NW_OK;
    const unsigned grid = (unsigned)((M + 255) / 256);
    hipStream_t s = (hipStream_t)stream;
#define NW_LTI_ND(K_)                                                                              \
    case K_:                                                                                       \
        if (num_inputs == 1)                                                                       \
            k_lti_nd<K_, 1><<<grid, 256, 0, s>>>(q, node_table, M, states, state_stride,           \
                                                 conserve_sum, n_steps, weights, gmax, cond_out); \
        else                                                                                       \
            k_lti_nd<K_, 2><<<grid, 256, 0, s>>>(q, node_table, M, states, state_stride,           \
                                                 conserve_sum, n_steps, weights, gmax, cond_out); \
        break;
    switch (S) {
        NW_LTI_ND(1) NW_LTI_ND(2) NW_LTI_ND(3) NW_LTI_ND(4) NW_LTI_ND(5) NW_LTI_ND(6)
        NW_LTI_ND(7) NW_LTI_ND(8) NW_LTI_ND(9) NW_LTI_ND(10) NW_LTI_ND(11) NW_LTI_ND(12)
        NW_LTI_ND(13) NW_LTI_ND(14) NW_LTI_ND(15) NW_LTI_ND(16)
        default: NW_REQUIRE(false, "state count %d not supported (1..16)", S);
    }
#undef NW_LTI_ND
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_ca_step(const nw_propagator* p, const nw_ca_params* prm, int32_t n_steps, float* ca,
               float* cab, float* dcdt, void* stream) {
    NW_REQUIRE(p && prm && ca && cab && dcdt, "nw_ca_step: null argument");
    NW_REQUIRE(!p->symmetric && !p->chebyshev,
               "nw_ca_step takes the general format (not the symmetric or Chebyshev store)");
    NW_REQUIRE(n_steps >= 0 && prm->dt_ms > 0.0f, "bad step count or time step");
    if (n_steps == 0) return NW_OK;
    CaArgs a{p->cell_row0, p->cell_nz0, p->row_off, p->vals, p->cols, p->rowsum, p->order,
             p->max_cell_rows, n_steps, *prm, ca, cab, dcdt};
    const size_t lds = 2 * sizeof(float) * (size_t)p->max_cell_rows;
    NW_REQUIRE(lds <= 160 * 1024, "cell of %d segments exceeds LDS", p->max_cell_rows);
    const int bt = prm->block_threads ? prm->block_threads : 256;
    const unsigned grid = (unsigned)p->num_cells;
    hipStream_t s = (hipStream_t)stream;
    const bool lane_rows = p->max_row_nnz <= kLaneRowMax;
#define NW_CA_LAUNCH(B)                                                                   \
    case B:                                                                               \
        if (lane_rows) k_ca_step<B, 1><<<grid, B, lds, s>>>(a);                           \
        else k_ca_step<B, 0><<<grid, B, lds, s>>>(a);                                     \
        break;
    switch (bt) {
        NW_CA_LAUNCH(64) NW_CA_LAUNCH(128) NW_CA_LAUNCH(256) NW_CA_LAUNCH(512)
        default: NW_REQUIRE(false, "block_threads must be 64, 128, 256 or 512 (got %d)", bt);
    }
#undef NW_CA_LAUNCH
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_weighted_sum(int64_t n, const double* w, const float* x, const float* y, const float* z,
                    double z_scale, double* partials, int32_t num_partials, void* stream) {
    NW_REQUIRE(n >= 0 && w && x && partials && num_partials >= 1 && num_partials <= 65536,
               "nw_weighted_sum: bad argument");
    k_weighted_sum<<<(unsigned)num_partials, 256, 0, (hipStream_t)stream>>>(n, w, x, y, z, z_scale,
                                                                            partials);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_weighted_sum_total(int64_t n, const double* w, const float* x, const float* y,
                          const float* z, double z_scale, int32_t num_partials, double* total,
                          void* stream) {
    NW_REQUIRE(n >= 0 && w && x && total, "nw_weighted_sum_total: null argument");
    NW_REQUIRE(num_partials >= 1 && num_partials <= kSumScratch,
               "nw_weighted_sum_total: num_partials %d outside 1..%d (the partials array it holds "
               "per device)", num_partials, kSumScratch);
    double* part = nullptr;
    int rc = scratch_of_device(&part);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    // every partial the sum reads is written by this call's first kernel: nothing to clear
    k_weighted_sum<<<(unsigned)num_partials, 256, 0, s>>>(n, w, x, y, z, z_scale, part);
    NW_HIP(hipGetLastError());
    k_sum_in_order<<<1, 64, 0, s>>>(part, num_partials, total);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_ap_detect(int64_t num, const int32_t* seg, const float* v, float threshold,
                 uint8_t* state, uint8_t* detected, void* stream) {
    NW_REQUIRE(num >= 0 && (num == 0 || (seg && v && state && detected)), "nw_ap_detect: null argument");
    if (num == 0) return NW_OK;
    k_ap_detect<<<(unsigned)((num + 255) / 256), 256, 0, (hipStream_t)stream>>>(num, seg, v, threshold,
                                                                              state, detected);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

int nw_accumulate(int64_t num_segments, const int64_t* seg_inst_ptr, const float* g_inst,
                  const int64_t* g_index, const float* e_inst, const float* nonspec_i, float* sum_g,
                  float* drive_v, float* sum_i, void* stream) {
    NW_REQUIRE(num_segments >= 0 && seg_inst_ptr && sum_g && drive_v && sum_i,
               "nw_accumulate: null argument");
    if (num_segments == 0) return NW_OK;
    k_accumulate<<<(unsigned)((num_segments + kAccBlock - 1) / kAccBlock), kAccBlock, 0,
                   (hipStream_t)stream>>>(num_segments, seg_inst_ptr, g_inst, g_index, e_inst,
                                          nonspec_i, sum_g, drive_v, sum_i);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

}  // extern "C"
