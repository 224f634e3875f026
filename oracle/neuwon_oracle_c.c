/*
 * FP64 C restatement of the staggered HH advance -- TEST INFRASTRUCTURE ONLY (the checker and the
 * timed CPU baseline; the product never links it).
 *
 * Same algorithm as oracle/neuwon_oracle.py:HHSim, step for step:
 *   half step (rxd_model.py:91-113, electric.py:57-76):
 *       d = nan_to_num(d_num / sum_g); x = d - (d - v) exp(-dt_s sum_g / C) + dt I / C; v = P x
 *   mechanisms (rxd_model.py:115-123): every gate a two-state kinetic scheme advanced by the
 *       generated LTI kernel (lti_sim/codegen.py:39-106, polynomial table [bucket][col][row][term],
 *       CONSERVE renormalisation) or, in gate mode 1, the exact hh.mod update (solver.py:48-55);
 *       g = gbar m^3 h + gbar n^4 + gbar_l, d_num = sum g E (hh.mod:41-46, species.py:256-260)
 *   half step again.
 * Cells are independent, so threads take whole cells (OpenMP, dynamic schedule).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double dt, vmin, vmax;
    int32_t num_buckets;
    double e_na, e_k, e_l;
    int32_t stim_mode; /* 0 none, 1 constant, 2 Poisson pulses (at segment 0 of every cell) */
    double stim_amp;   /* A */
    uint32_t seed, threshold;
    int32_t pulse_ticks;
    int32_t gate_mode; /* 0 tables, 1 exact */
    double celsius;
    /* FP32-storage emulation (not part of the reference's algorithm: it measures the error floor
     * that storing the state in FP32 sets): bit 0 rounds every gate to FP32 after its update,
     * bit 1 every voltage to FP32 at the end of every step */
    int32_t fp32_state;
} oracle_params;

static uint32_t stim_hash(uint32_t seed, uint32_t cell, uint32_t tick) {
    uint32_t h = (seed * 0x9E3779B1u) ^ (cell * 0x85EBCA77u) ^ (tick * 0xC2B2AE3Du);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

uint32_t oracle_stim_hash(uint32_t seed, uint32_t cell, uint32_t tick) {
    return stim_hash(seed, cell, tick);
}

static double stim_current(const oracle_params* p, uint32_t cell, int64_t tick) {
    if (p->stim_mode == 1) return p->stim_amp;
    if (p->stim_mode == 2) {
        int k = 0;
        for (int j = 0; j < p->pulse_ticks; ++j) {
            int64_t t = tick - j;
            if (t < 0) break;
            if (stim_hash(p->seed, cell, (uint32_t)t) < p->threshold) ++k;
        }
        return p->stim_amp * k;
    }
    return 0.0;
}

/* hh.mod:88-94 */
static double vtrap(double x, double y) {
    if (fabs(x / y) < 1e-6) return y * (1.0 - x / y / 2.0);
    return x / (exp(x / y) - 1.0);
}

static void rates(int gate, double v, double q10, double* a, double* b) {
    switch (gate) {
        case 0: *a = 0.1 * vtrap(-(v + 40.0), 10.0); *b = 4.0 * exp(-(v + 65.0) / 18.0); break;
        case 1: *a = 0.07 * exp(-(v + 65.0) / 20.0); *b = 1.0 / (exp(-(v + 35.0) / 10.0) + 1.0); break;
        default: *a = 0.01 * vtrap(-(v + 55.0), 10.0); *b = 0.125 * exp(-(v + 65.0) / 80.0); break;
    }
    *a *= q10;
    *b *= q10;
}

/* codegen.py:39-106 for a two-state scheme with CONSERVE sum 1; poly = [bucket][col][row][term]. */
static double lti_two_state(const double* poly, const oracle_params* p, double v, double open) {
    const int nb = p->num_buckets;
    const double frq = nb / (p->vmax - p->vmin);
    double u = (v - p->vmin) * frq;
    int64_t b = (int64_t)u; /* truncation toward zero */
    double f;
    if (b > nb - 1) { b = nb - 1; f = 1.0; }
    else if (b < 0) { b = 0; f = 0.0; }
    else f = u - (double)b;
    const double* t = poly + b * 8;
    double s[2] = {1.0 - open, open};
    double scratch[2] = {0.0, 0.0};
    for (int col = 0; col < 2; ++col)
        for (int row = 0; row < 2; ++row) {
            const double* c = t + (col * 2 + row) * 2;
            double poly_v = 0.0;
            poly_v += c[0];
            poly_v += f * c[1];
            scratch[row] += poly_v * s[col];
        }
    double corr = 1.0 / (scratch[0] + scratch[1]);
    return scratch[1] * corr;
}

static void half_step(const oracle_params* p, int64_t r0, int64_t n, const int64_t* rowptr,
                      const int32_t* col, const double* val, const double* C, double* v,
                      const double* sum_g, const double* d_num, uint32_t cell, int64_t tick,
                      double* x) {
    const double half = 0.5 * p->dt, dt_s = half * 1e-3;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t gi = r0 + i;
        double d = d_num[gi] / sum_g[gi];
        if (!isfinite(d)) d = 0.0; /* nan_to_num, rxd_model.py:109 */
        const double I = (i == 0) ? stim_current(p, cell, tick) : 0.0;
        const double dv = half * I / C[gi];
        const double expo = -dt_s * sum_g[gi] / C[gi];
        const double alpha = exp(expo);
        const double diff = d - v[gi];
        x[i] = d - diff * alpha + dv;
    }
    for (int64_t i = 0; i < n; ++i) {
        const int64_t gi = r0 + i;
        double acc = 0.0;
        for (int64_t k = rowptr[gi]; k < rowptr[gi + 1]; ++k) acc += val[k] * x[col[k] - r0];
        v[gi] = acc;
    }
}

/* Advance every cell by n_steps; returns 0. */
int oracle_hh_run(const oracle_params* p, int64_t num_cells, const int64_t* cell_ptr,
                  const int64_t* rowptr, const int32_t* col, const double* val, const double* C,
                  const double* gna, const double* gk, const double* gl, double* v, double* m,
                  double* h, double* n, double* sum_g, double* d_num, const double* poly,
                  int64_t step0, int32_t n_steps, int32_t threads) {
    int64_t nmax = 0;
    for (int64_t c = 0; c < num_cells; ++c)
        if (cell_ptr[c + 1] - cell_ptr[c] > nmax) nmax = cell_ptr[c + 1] - cell_ptr[c];
    const double q10 = pow(3.0, (p->celsius - 6.3) / 10.0);
    double* gates[3] = {m, h, n};
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
#endif
    {
        double* x = (double*)malloc(sizeof(double) * (size_t)(nmax > 0 ? nmax : 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t c = 0; c < num_cells; ++c) {
            const int64_t r0 = cell_ptr[c], cn = cell_ptr[c + 1] - r0;
            for (int32_t s = 0; s < n_steps; ++s) {
                const int64_t tick = 2 * (step0 + s) + 1;
                half_step(p, r0, cn, rowptr, col, val, C, v, sum_g, d_num, (uint32_t)c, tick, x);
                for (int64_t i = 0; i < cn; ++i) {
                    const int64_t gi = r0 + i;
                    for (int g = 0; g < 3; ++g) {
                        double* st = gates[g];
                        if (p->gate_mode == 1) {
                            double a, b;
                            rates(g, v[gi], q10, &a, &b);
                            double sm = a + b, tau = 1.0 / sm, inf = a / sm;
                            double c1 = -1.0 / tau, c0 = inf / tau;
                            st[gi] = -c0 / c1 + (c0 + c1 * st[gi]) * exp(c1 * p->dt) / c1;
                        } else {
                            st[gi] = lti_two_state(poly + (int64_t)g * p->num_buckets * 8, p, v[gi],
                                                   st[gi]);
                        }
                        if (p->fp32_state & 1) st[gi] = (double)(float)st[gi];
                    }
                    const double mm = m[gi], hh = h[gi], nn = n[gi];
                    const double g_na = gna[gi] * mm * mm * mm * hh;
                    const double g_k = gk[gi] * nn * nn * nn * nn;
                    const double g_l = gl[gi];
                    sum_g[gi] = g_na + g_k + g_l;
                    d_num[gi] = g_na * p->e_na + g_k * p->e_k + g_l * p->e_l;
                }
                half_step(p, r0, cn, rowptr, col, val, C, v, sum_g, d_num, (uint32_t)c, tick + 1, x);
                if (p->fp32_state & 2)
                    for (int64_t i = 0; i < cn; ++i) v[r0 + i] = (double)(float)v[r0 + i];
            }
        }
        free(x);
    }
    return 0;
}

/* The generated LTI kernel (lti_sim/codegen.py:39-106) for M instances of an S-state scheme with
 * one linear input: poly = [bucket][col][row][term] (num_buckets buckets over vmin..vmax), states
 * [S][M] in place.  Step k reads the voltage v0[i] + wave[k] (wave may be NULL: held).  CONSERVE
 * renormalisation when conserve > 0.  Instances are independent: OpenMP over instances. */
int oracle_lti_run(int32_t S, int32_t num_buckets, int32_t terms, double vmin, double vmax,
                   const double* poly, int64_t M, double* states, const double* v0,
                   const double* wave, int32_t n_steps, double conserve, int32_t threads) {
    if (S < 1 || S > 32 || terms < 1 || terms > 8) return -1;
    const double frq = num_buckets / (vmax - vmin);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < M; ++i) {
        double s[32], scratch[32], basis[8];
        for (int k = 0; k < S; ++k) s[k] = states[(int64_t)k * M + i];
        for (int32_t step = 0; step < n_steps; ++step) {
            const double v = v0[i] + (wave ? wave[step] : 0.0);
            const double u = (v - vmin) * frq;
            int64_t b = (int64_t)u; /* truncation toward zero */
            double f;
            if (b > num_buckets - 1) { b = num_buckets - 1; f = 1.0; }
            else if (b < 0) { b = 0; f = 0.0; }
            else f = u - (double)b;
            basis[0] = 1.0;
            for (int t = 1; t < terms; ++t) basis[t] = basis[t - 1] * f;
            for (int r = 0; r < S; ++r) scratch[r] = 0.0;
            const double* tb = poly + b * (int64_t)S * S * terms;
            for (int col = 0; col < S; ++col)
                for (int row = 0; row < S; ++row) {
                    const double* c = tb + ((int64_t)col * S + row) * terms;
                    double p = 0.0;
                    for (int t = 0; t < terms; ++t) p += c[t] * basis[t];
                    scratch[row] += p * s[col];
                }
            if (conserve > 0.0) {
                double tot = 0.0;
                for (int r = 0; r < S; ++r) tot += scratch[r];
                const double corr = conserve / tot;
                for (int r = 0; r < S; ++r) scratch[r] *= corr;
            }
            for (int r = 0; r < S; ++r) s[r] = scratch[r];
        }
        for (int k = 0; k < S; ++k) states[(int64_t)k * M + i] = s[k];
    }
    return 0;
}
