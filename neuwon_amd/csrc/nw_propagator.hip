// Exact-integration electric propagator: FP64 construction on the device and the CSR half step.
//
// Construction replaces Electric._compute_propagator_matrix (neuwon/rxd/neuron/electric.py:79-117),
// which assembles A*dt from the segment tree and calls scipy.sparse.linalg.expm on the whole,
// block-diagonal matrix.  Here every cell block is exponentiated on its own, densely, in FP64:
//
//   M = A*dt has zero row sums and non-negative off-diagonals.  With lambda = max |M_ii| the
//   shifted matrix B = M + lambda*I is entrywise non-negative with row sums exactly lambda, so
//     expm(M) = (exp(-lambda/2^s) * expm(B/2^s))^(2^s)
//   where lambda/2^s <= 1/2.  expm(B/2^s) is a degree-15 Taylor polynomial evaluated by the
//   Paterson-Stockmeyer scheme (6 GEMMs); s squarings follow.  Every term and every product is
//   non-negative, so no cancellation occurs and tiny entries keep full relative accuracy -- the
//   pruning threshold is applied to accurate values.  GEMMs run on rocBLAS dgemm (FP64 matrix
//   cores); the remaining steps are the kernels below.
//
// The half step replaces Electric._advance_electric (electric.py:57-76).  One workgroup owns one
// cell: the relaxed vector x = d - (d - v)*alpha + dt*I/C is staged in LDS, then every wavefront
// takes whole rows of the cell's CSR block (rows hold hundreds of entries), streams the FP32 values
// and 16-bit local columns with coalesced loads, gathers x from LDS and reduces across its lanes.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "nw_internal.h"

namespace nw {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

#define NW_BLAS(expr)                                                                      \
    do {                                                                                   \
        rocblas_status s_ = (expr);                                                        \
        if (s_ != rocblas_status_success) {                                                \
            ::nw::set_error("%s failed: rocblas status %d (%s:%d)", #expr, (int)s_,        \
                            __FILE__, __LINE__);                                           \
            return NW_ERR_BLAS;                                                            \
        }                                                                                  \
    } while (0)

// ----------------------------------------------------------------------------------------------
// Construction kernels (row-major n x n FP64 matrices)
// ----------------------------------------------------------------------------------------------
__global__ void k_scatter(double* X, int n, int m, const int32_t* ri, const int32_t* ci,
                          const double* vals) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) X[(int64_t)ri[k] * n + ci[k]] = vals[k];
}

__global__ void k_set_diag(double* X, int n, const double* d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) X[(int64_t)i * n + i] = d[i];
}

// out = a0*I + a1*X1 + a2*X2 + a3*X3
__global__ void k_lincomb(double* out, const double* X1, const double* X2, const double* X3,
                          int n, double a0, double a1, double a2, double a3) {
    int64_t total = (int64_t)n * n;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        double v = a1 * X1[k] + a2 * X2[k] + a3 * X3[k];
        int64_t i = k / n, j = k - i * n;
        if (i == j) v += a0;
        out[k] = v;
    }
}

// One wavefront per row: number of entries kept (diagonal always kept, off-diagonal kept when
// |P| >= prune, electric.py:114).
__global__ void k_count(const double* P, int n, double prune, int32_t* counts) {
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    const double* pr = P + (int64_t)row * n;
    int cnt = 0;
    for (int j0 = 0; j0 < n; j0 += kWave) {
        int j = j0 + lane;
        bool keep = (j < n) && (j == row || fabs(pr[j]) >= prune);
        cnt += __popcll(__ballot(keep));
    }
    if (lane == 0) counts[row] = cnt;
}

// One wavefront per row: store kept entries in column order.  FP32 variant: the diagonal is set so
// that the FP32 row sums (in FP64) to the FP64 sum of the kept entries.
template <bool kF32>
__global__ void k_fill(const double* P, int n, double prune, const uint32_t* row_off_local,
                       float* vals32, uint16_t* cols16, double* rowsum, double* vals64,
                       int32_t* cols32) {
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    const double* pr = P + (int64_t)row * n;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    float diag32 = 0.0f;
    if (kF32) {
        double s64 = 0.0, s32 = 0.0;
        for (int j0 = 0; j0 < n; j0 += kWave) {
            int j = j0 + lane;
            if (j < n) {
                double p = pr[j];
                if (j == row || fabs(p) >= prune) {
                    s64 += p;
                    if (j != row) s32 += (double)(float)p;
                }
            }
        }
        s64 = wave_sum(s64);
        s32 = wave_sum(s32);
        diag32 = (float)(s64 - s32);
        if (lane == 0) rowsum[row] = s32 + (double)diag32;
    }
    const uint32_t beg = row_off_local[row], len = row_off_local[row + 1] - beg;
    uint32_t pos = 0;  // entries of the row written so far (column order)
    for (int j0 = 0; j0 < n; j0 += kWave) {
        int j = j0 + lane;
        double p = (j < n) ? pr[j] : 0.0;
        bool keep = (j < n) && (j == row || fabs(p) >= prune);
        uint64_t mask = __ballot(keep);
        if (keep) {
            uint32_t k = pos + (uint32_t)__popcll(mask & lt);
            if (kF32) {
                const uint32_t at = beg + stored_slot(k, len);
                vals32[at] = (j == row) ? diag32 : (float)p;
                cols16[at] = (uint16_t)j;
            } else {
                vals64[beg + k] = p;
                cols32[beg + k] = j;
            }
        }
        pos += (uint32_t)__popcll(mask);
    }
    if (kF32) {  // padding up to the next row: value 0, column = the row itself
        for (uint32_t k = pos + lane; k < len; k += kWave) {
            const uint32_t at = beg + stored_slot(k, len);
            vals32[at] = 0.0f;
            cols16[at] = (uint16_t)row;
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Symmetric format.  S = W P W^-1 with w = sqrt(C) is symmetric (A = C^-1 L, L symmetric); the
// two triangles of the computed P are averaged into S so S_ij and S_ji are bitwise equal.
// A pair (i, j) is kept when max(P_ij, P_ji) >= prune (the reference prunes P entry by entry,
// electric.py:114; one triangle can only store the pair, so the pair is kept if either entry is).
// ----------------------------------------------------------------------------------------------
__global__ void k_symmetrize(const double* __restrict__ P, const double* __restrict__ w, int n,
                             double* __restrict__ S) {
    __shared__ double tile[32][33];
    const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {             // tile of P at (bj, bi), transposed into LDS
        const int r = bj + k, c = bi + tx;
        tile[tx][k] = (r < n && c < n) ? P[(int64_t)r * n + c] : 0.0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int i = bi + k, j = bj + tx;
        if (i < n && j < n) {
            const double pij = P[(int64_t)i * n + j], pji = tile[k][tx];
            S[(int64_t)i * n + j] = 0.5 * ((w[i] / w[j]) * pij + (w[j] / w[i]) * pji);
        }
    }
}

__device__ __forceinline__ bool keep_pair(double sij, double wi, double wj, double prune) {
    const double r = wj / wi;
    return fabs(sij) * fmax(r, 1.0 / r) >= prune;
}

__global__ void k_count_sym(const double* S, const double* w, int n, double prune, int32_t* counts) {
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    const double* sr = S + (int64_t)row * n;
    const double wi = w[row];
    int cnt = 1;  // the diagonal
    for (int j0 = row + 1; j0 < n; j0 += kWave) {
        const int j = j0 + lane;
        const bool keep = (j < n) && keep_pair(sr[j], wi, w[j], prune);
        cnt += __popcll(__ballot(keep));
    }
    if (lane == 0) counts[row] = cnt;
}

// Row `row` of the symmetric store: the diagonal, then kept entries j > row in column order
// (placed by stored_slot).  The
// FP32 diagonal is chosen so that the implied FP32 row of P, sum_j S32_ij w_j / w_i over the whole
// row (both triangles), sums in FP64 to the FP64 row sum of P over the kept pairs.
__global__ void k_fill_sym(const double* S, const double* w, int n, double prune,
                           const uint32_t* row_off_local, float* vals32, uint16_t* cols16,
                           double* rowsum, double* wgt) {
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    const double* sr = S + (int64_t)row * n;
    const double wi = w[row];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    double s64 = 0.0, s32 = 0.0;
    for (int j0 = 0; j0 < n; j0 += kWave) {
        const int j = j0 + lane;
        if (j < n && j != row && keep_pair(sr[j], wi, w[j], prune)) {
            const double r = w[j] / wi;
            s64 += sr[j] * r;
            s32 += (double)(float)sr[j] * r;
        }
    }
    s64 = wave_sum(s64) + sr[row];
    s32 = wave_sum(s32);
    const float diag32 = (float)(s64 - s32);
    const uint32_t beg = row_off_local[row], len = row_off_local[row + 1] - beg;
    if (lane == 0) {
        rowsum[row] = s32 + (double)diag32;
        wgt[row] = wi;
        vals32[beg + stored_slot(0, len)] = diag32;
        cols16[beg + stored_slot(0, len)] = (uint16_t)row;
    }
    uint32_t pos = 1;  // entries of the row written so far (column order)
    for (int j0 = row + 1; j0 < n; j0 += kWave) {
        const int j = j0 + lane;
        const bool keep = (j < n) && keep_pair(sr[j], wi, w[j], prune);
        const uint64_t mask = __ballot(keep);
        if (keep) {
            const uint32_t at = beg + stored_slot(pos + (uint32_t)__popcll(mask & lt), len);
            vals32[at] = (float)sr[j];
            cols16[at] = (uint16_t)j;
        }
        pos += (uint32_t)__popcll(mask);
    }
    for (uint32_t k = pos + lane; k < len; k += kWave) {
        const uint32_t at = beg + stored_slot(k, len);
        vals32[at] = 0.0f;
        cols16[at] = (uint16_t)row;
    }
}

// ----------------------------------------------------------------------------------------------
// Host-side construction driver
// ----------------------------------------------------------------------------------------------
struct Workspace {
    int nmax = 0;
    double* X[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int32_t *d_ri = nullptr, *d_ci = nullptr, *d_counts = nullptr;
    double *d_v = nullptr, *d_diag = nullptr, *d_w = nullptr;
    uint32_t* d_off = nullptr;
    rocblas_handle blas = nullptr;
    hipStream_t stream = nullptr;
    // the series path's per-cell arrays (grown on demand)
    int32_t *s_nbi = nullptr, *s_ovi = nullptr;
    double *s_nbc = nullptr, *s_ovc = nullptr, *s_dg = nullptr, *s_c = nullptr, *s_C = nullptr;
    int s_rows = 0, s_ov = 0, s_terms = 0;

    // dense: the Taylor path's six n x n buffers (cells above kSeriesMaxRows rows); otherwise
    // two (P and its symmetrised copy)
    int init(int n, bool dense = true) {
        nmax = n;
        size_t nn = (size_t)n * n * sizeof(double);
        for (int i = 0; i < (dense ? 6 : 2); ++i) NW_HIP(hipMalloc(&X[i], nn));
        NW_HIP(hipMalloc(&d_ri, sizeof(int32_t) * 2 * n));
        NW_HIP(hipMalloc(&d_ci, sizeof(int32_t) * 2 * n));
        NW_HIP(hipMalloc(&d_v, sizeof(double) * 2 * n));
        NW_HIP(hipMalloc(&d_diag, sizeof(double) * n));
        NW_HIP(hipMalloc(&d_w, sizeof(double) * n));
        NW_HIP(hipMalloc(&d_counts, sizeof(int32_t) * n));
        NW_HIP(hipMalloc(&d_off, sizeof(uint32_t) * (n + 1)));
        NW_HIP(hipStreamCreate(&stream));
        NW_BLAS(rocblas_create_handle(&blas));
        NW_BLAS(rocblas_set_stream(blas, stream));
        NW_BLAS(rocblas_set_pointer_mode(blas, rocblas_pointer_mode_host));
        return NW_OK;
    }
    int series_arrays(int n, int ov, int terms) {
        if (n > s_rows) {
            if (s_nbi) (void)hipFree(s_nbi);
            if (s_nbc) (void)hipFree(s_nbc);
            if (s_dg) (void)hipFree(s_dg);
            if (s_C) (void)hipFree(s_C);
            s_nbi = nullptr;
            s_nbc = s_dg = s_C = nullptr;
            NW_HIP(hipMalloc(&s_nbi, sizeof(int32_t) * 3 * n));
            NW_HIP(hipMalloc(&s_nbc, sizeof(double) * 3 * n));
            NW_HIP(hipMalloc(&s_dg, sizeof(double) * n));
            NW_HIP(hipMalloc(&s_C, sizeof(double) * n));
            s_rows = n;
        }
        if (ov > s_ov) {
            if (s_ovi) (void)hipFree(s_ovi);
            if (s_ovc) (void)hipFree(s_ovc);
            s_ovi = nullptr;
            s_ovc = nullptr;
            NW_HIP(hipMalloc(&s_ovi, sizeof(int32_t) * ov));
            NW_HIP(hipMalloc(&s_ovc, sizeof(double) * ov));
            s_ov = ov;
        }
        if (terms > s_terms) {
            if (s_c) (void)hipFree(s_c);
            s_c = nullptr;
            NW_HIP(hipMalloc(&s_c, sizeof(double) * terms));
            s_terms = terms;
        }
        return NW_OK;
    }
    ~Workspace() {
        for (int32_t* x : {s_nbi, s_ovi})
            if (x) (void)hipFree(x);
        for (double* x : {s_nbc, s_ovc, s_dg, s_c, s_C})
            if (x) (void)hipFree(x);
        for (auto& x : X)
            if (x) (void)hipFree(x);
        if (d_ri) (void)hipFree(d_ri);
        if (d_ci) (void)hipFree(d_ci);
        if (d_v) (void)hipFree(d_v);
        if (d_diag) (void)hipFree(d_diag);
        if (d_w) (void)hipFree(d_w);
        if (d_counts) (void)hipFree(d_counts);
        if (d_off) (void)hipFree(d_off);
        if (blas) rocblas_destroy_handle(blas);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int gemm(Workspace& w, int n, const double* A, const double* B, double* C, double beta) {
    const double one = 1.0;
    NW_BLAS(rocblas_dgemm(w.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                          A, n, B, n, &beta, C, n));
    return NW_OK;
}

// Dense FP64 expm of one cell's A*dt; on success *result points at one of the workspace buffers.
static int expm_cell(Workspace& w, int n, const int32_t* parent, const double* R, const double* C,
                     double dt_ms, double** result) {
    const double dt = dt_ms * 1e-3;  // seconds, electric.py:86
    std::vector<double> diag(n, 0.0);
    std::vector<int32_t> ri, ci;
    std::vector<double> vv;
    ri.reserve(2 * n);
    ci.reserve(2 * n);
    vv.reserve(2 * n);
    for (int c = 0; c < n; ++c) {  // electric.py:92-108
        int p = parent[c];
        if (p < 0) continue;
        NW_REQUIRE(p < n, "parent index %d out of range for cell of %d segments", p, n);
        double r = R[c];
        ri.push_back(p); ci.push_back(c); vv.push_back(+dt / (r * C[p]));
        diag[c] -= dt / (r * C[c]);
        ri.push_back(c); ci.push_back(p); vv.push_back(+dt / (r * C[c]));
        diag[p] -= dt / (r * C[p]);
    }
    double lambda = 0.0;
    for (double d : diag) lambda = std::max(lambda, -d);
    int s = 0;
    while (lambda * std::ldexp(1.0, -s) > 0.5) ++s;
    const double scale = std::ldexp(1.0, -s);
    const double mu = lambda * scale;
    for (auto& x : vv) x *= scale;
    for (int i = 0; i < n; ++i) diag[i] = (diag[i] + lambda) * scale;
    const int m = (int)vv.size();

    double *X1 = w.X[0], *X2 = w.X[1], *X3 = w.X[2], *X4 = w.X[3], *T = w.X[4], *E = w.X[5];
    const size_t nn = (size_t)n * n;
    NW_HIP(hipMemsetAsync(X1, 0, nn * sizeof(double), w.stream));
    if (m > 0) {
        NW_HIP(hipMemcpyAsync(w.d_ri, ri.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, w.stream));
        NW_HIP(hipMemcpyAsync(w.d_ci, ci.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, w.stream));
        NW_HIP(hipMemcpyAsync(w.d_v, vv.data(), m * sizeof(double), hipMemcpyHostToDevice, w.stream));
        k_scatter<<<(m + 255) / 256, 256, 0, w.stream>>>(X1, n, m, w.d_ri, w.d_ci, w.d_v);
    }
    NW_HIP(hipMemcpyAsync(w.d_diag, diag.data(), n * sizeof(double), hipMemcpyHostToDevice, w.stream));
    k_set_diag<<<(n + 255) / 256, 256, 0, w.stream>>>(X1, n, w.d_diag);
    NW_HIP(hipGetLastError());

    // Taylor coefficients exp(-mu)/k!, k = 0..15
    double c[16];
    c[0] = std::exp(-mu);
    for (int k = 1; k < 16; ++k) c[k] = c[k - 1] / k;
    int rc;
    if ((rc = gemm(w, n, X1, X1, X2, 0.0))) return rc;
    if ((rc = gemm(w, n, X2, X1, X3, 0.0))) return rc;
    if ((rc = gemm(w, n, X2, X2, X4, 0.0))) return rc;
    const int lb = 256;
    const int lg = (int)std::min<int64_t>(((int64_t)nn + lb - 1) / lb, 8192);
    // Paterson-Stockmeyer: E = Q0 + X4 (Q1 + X4 (Q2 + X4 Q3)),  Qj = sum_{i<4} c_{4j+i} X^i
    k_lincomb<<<lg, lb, 0, w.stream>>>(T, X1, X2, X3, n, c[12], c[13], c[14], c[15]);
    k_lincomb<<<lg, lb, 0, w.stream>>>(E, X1, X2, X3, n, c[8], c[9], c[10], c[11]);
    if ((rc = gemm(w, n, X4, T, E, 1.0))) return rc;
    k_lincomb<<<lg, lb, 0, w.stream>>>(T, X1, X2, X3, n, c[4], c[5], c[6], c[7]);
    if ((rc = gemm(w, n, X4, E, T, 1.0))) return rc;
    k_lincomb<<<lg, lb, 0, w.stream>>>(E, X1, X2, X3, n, c[0], c[1], c[2], c[3]);
    if ((rc = gemm(w, n, X4, T, E, 1.0))) return rc;
    NW_HIP(hipGetLastError());
    double* cur = E;
    double* nxt = T;
    for (int k = 0; k < s; ++k) {
        if ((rc = gemm(w, n, cur, cur, nxt, 0.0))) return rc;
        std::swap(cur, nxt);
    }
    *result = cur;
    return NW_OK;
}

// Row counts -> local CSR offsets on the host (each row rounded up to `align` entries).
static int count_rows(Workspace& w, const double* P, int n, double prune, int align,
                      std::vector<uint32_t>& off, int32_t& max_row, int64_t& kept,
                      bool sym = false) {
    const int rows_per_block = 4;
    if (sym)
        k_count_sym<<<(n + rows_per_block - 1) / rows_per_block, rows_per_block * kWave, 0,
                      w.stream>>>(P, w.d_w, n, prune, w.d_counts);
    else
        k_count<<<(n + rows_per_block - 1) / rows_per_block, rows_per_block * kWave, 0,
                  w.stream>>>(P, n, prune, w.d_counts);
    NW_HIP(hipGetLastError());
    std::vector<int32_t> counts(n);
    NW_HIP(hipMemcpyAsync(counts.data(), w.d_counts, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                          w.stream));
    NW_HIP(hipStreamSynchronize(w.stream));
    off.assign(n + 1, 0);
    max_row = 0;
    kept = 0;
    uint64_t acc = 0;
    for (int i = 0; i < n; ++i) {
        off[i] = (uint32_t)acc;
        acc += (uint64_t)((counts[i] + align - 1) / align * align);
        kept += sym ? 2 * (int64_t)counts[i] - 1 : counts[i];
        max_row = std::max(max_row, counts[i]);
    }
    NW_REQUIRE(acc < 0xFFFFFFFFull, "cell block too large for 32-bit row offsets");
    off[n] = (uint32_t)acc;
    return NW_OK;
}

static int grow(nw_propagator* p, int64_t need) {
    if (need <= p->vals_capacity) return NW_OK;
    int64_t cap = std::max<int64_t>(need, (int64_t)(p->vals_capacity * 1.5));
    float* nv = nullptr;
    uint16_t* nc = nullptr;
    if (hipMalloc(&nv, cap * sizeof(float)) != hipSuccess ||
        hipMalloc(&nc, cap * sizeof(uint16_t)) != hipSuccess) {
        if (nv) (void)hipFree(nv);
        set_error("out of device memory growing the propagator to %lld entries", (long long)cap);
        return NW_ERR_ALLOC;
    }
    if (p->stored > 0) {
        NW_HIP(hipMemcpy(nv, p->vals, p->stored * sizeof(float), hipMemcpyDeviceToDevice));
        NW_HIP(hipMemcpy(nc, p->cols, p->stored * sizeof(uint16_t), hipMemcpyDeviceToDevice));
    }
    if (p->vals) (void)hipFree(p->vals);
    if (p->cols) (void)hipFree(p->cols);
    p->vals = nv;
    p->cols = nc;
    p->vals_capacity = cap;
    return NW_OK;
}

// c_0 = e^-y I_0(y), c_k = 2 e^-y I_k(y) (y = rho/2), k = 0..K, K the last index with c_k > tol,
// the dropped terms' sum added to c_0.
// Miller's backward recurrence I_{k-1} = (2k/y) I_k + I_{k+1} from an index where I_k/I_0 is far
// below tol (I_k/I_0 ~ exp(-k^2 / 2y)), rescaled against overflow, normalised with
// e^y = I_0(y) + 2 sum_k I_k(y).
static int exp_chebyshev(double rho, double tol, std::vector<double>& c) {
    c.clear();
    const double y = 0.5 * rho;
    if (!(y > 1e-300)) {
        c.push_back(1.0);
        return 0;
    }
    const int M = (int)std::ceil(std::sqrt(2.0 * y * 120.0)) + 40;
    std::vector<double> J(M + 2, 0.0);
    J[M] = 1e-280;
    for (int k = M; k >= 1; --k) {
        J[k - 1] = (2.0 * k / y) * J[k] + J[k + 1];
        if (J[k - 1] > 1e250)
            for (int q = k - 1; q <= M; ++q) J[q] *= 1e-250;
    }
    double S = J[0];
    for (int k = 1; k <= M; ++k) S += 2.0 * J[k];
    int K = 0;
    for (int k = 1; k <= M; ++k)
        if (2.0 * J[k] / S > tol) K = k;
    if (K >= M - 10) return -1;  // the start index was too small
    c.resize(K + 1);
    c[0] = J[0] / S;
    double tail = 0.0;
    for (int k = K + 1; k <= M; ++k) tail += 2.0 * J[k] / S;
    for (int k = 1; k <= K; ++k) c[k] = 2.0 * J[k] / S;
    // the dropped tail is added to c_0: the operator keeps constant vectors (T_k(A~) 1 = 1), so
    // the truncated series then keeps them exactly as exp(A dt) does (P 1 = 1)
    c[0] += tail;
    return 0;
}

// Number of eigenvalues below sigma of the symmetric tree matrix M (diagonal dg, entry e[i] between
// segment i and its parent, parents before children): the negative pivots of M - sigma I
// eliminated from the leaves up (Sylvester's law of inertia).
static int count_below(int n, const int32_t* parent, const double* dg, const double* e,
                       double sigma, std::vector<double>& d) {
    for (int i = 0; i < n; ++i) d[i] = dg[i] - sigma;
    int neg = 0;
    for (int i = n - 1; i >= 0; --i) {
        double di = d[i];
        if (di == 0.0) di = -1e-300 * (std::abs(dg[i]) + 1.0);
        if (di < 0.0) ++neg;
        if (parent[i] >= 0) d[parent[i]] -= e[i] * e[i] / di;
    }
    return neg;
}

// Spectral radius of -B (B = A dt, similar to the symmetric M = -C^1/2 B C^-1/2, which is
// positive semidefinite): bisection between max_i M_ii (a Rayleigh quotient, below it) and the
// Gershgorin bound 2 max_i M_ii (above it) with inertia counts, returned with a 1e-8 relative
// margin above the last bound found to hold every eigenvalue.
static double tree_spectral_radius(int n, const int32_t* parent, const double* dg, const double* e,
                                   std::vector<double>& work) {
    double dmax = 0.0;
    for (int i = 0; i < n; ++i) dmax = std::max(dmax, dg[i]);
    if (!(dmax > 0.0)) return 0.0;
    double lo = dmax, hi = 2.0 * dmax * (1.0 + 1e-12);
    work.resize(n);
    if (count_below(n, parent, dg, e, hi, work) != n) return 2.0 * dmax * (1.0 + 1e-8);
    for (int it = 0; it < 45 && hi - lo > 1e-10 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (count_below(n, parent, dg, e, mid, work) == n) hi = mid;
        else lo = mid;
    }
    return hi * (1.0 + 1e-8);
}

// ----------------------------------------------------------------------------------------------
// P = exp(A dt) of one cell by uniformization (cells of at most kSeriesMaxRows rows).  With
// B = A dt (off-diagonal b_ij = dt / (R_edge C_i) >= 0, zero row sums) and sigma at least every
// |B_ii|, M = I + B / sigma is entrywise nonnegative with unit row sums, and
//   exp(B) = e^-sigma exp(sigma M) = sum_k w_k M^k,   w_k = e^-sigma sigma^k / k!
// (Poisson weights; Jensen's method).  Every term is nonnegative, so each entry of P, however
// small, is accurate to about K ulps relative (no cancellation, as in the nonnegative Taylor
// polynomial of the dense path), and the kept weights are normalised to sum to one, so P 1 = 1.
// Workgroup j forms column j, M^k e_j for k = 0 .. K, and writes row j of P from it: A = C^-1 G
// with G symmetric gives P^T = C P C^-1, so P_jr = C_r P_rj / C_j.  The work is about
// 2 (deg + 1) K n^2 flop per cell (K about sigma + 8 sqrt(sigma)) instead of the 30 n^3 of the
// dense Taylor path with its squarings.
// ----------------------------------------------------------------------------------------------
constexpr int kSeriesMaxRows = kChebRows * kChebBlock;
constexpr double kSeriesTol = 1e-17;      // Poisson mass dropped on each side
constexpr double kSeriesMargin = 1.02;    // sigma = margin * max_i |B_ii|: M_ii >= 0.02

// Poisson weights of mean sigma from the mode outwards (no underflow of e^-sigma), both tails
// dropped below tol of the total, normalised; returns the last index K (w[0 .. K]).
static int poisson_weights(double sigma, double tol, std::vector<double>& w) {
    w.clear();
    if (!(sigma > 0.0)) {
        w.push_back(1.0);
        return 0;
    }
    const int mode = (int)std::floor(sigma);
    std::vector<double> up(1, 1.0), down;
    for (int k = mode + 1;; ++k) {  // w_k / w_mode
        const double x = up.back() * sigma / k;
        if (x < 0.01 * tol && k > mode + 4) break;
        up.push_back(x);
    }
    for (int k = mode; k >= 1; --k) {
        const double x = (down.empty() ? 1.0 : down.back()) * k / sigma;
        if (x < 0.01 * tol) break;
        down.push_back(x);  // w_{k-1} / w_mode
    }
    const int lo = mode - (int)down.size();
    const int K = mode + (int)up.size() - 1;
    w.assign(K + 1, 0.0);
    for (size_t i = 0; i < down.size(); ++i) w[mode - 1 - i] = down[i];
    for (size_t i = 0; i < up.size(); ++i) w[mode + i] = up[i];
    double sum = 0.0;
    for (int k = K; k >= lo; --k) sum += w[k];
    // drop the tails below tol of the total, then renormalise
    double acc = 0.0;
    int last = K;
    while (last > mode && (acc += w[last]) < tol * sum) w[last--] = 0.0;
    acc = 0.0;
    for (int k = lo; k < mode && (acc += w[k]) < tol * sum; ++k) w[k] = 0.0;
    w.resize(last + 1);
    sum = 0.0;
    for (int k = last; k >= 0; --k) sum += w[k];
    for (auto& x : w) x /= sum;
    return last;
}

// NT threads, ROWS rows per thread (row r = threadIdx.x + NT q); nbi / nbc: three neighbour
// slots per row (local index, M_ij; an unused slot names the row itself with coefficient 0;
// slot 2 holds -(start + 1) of an overflow list ended by -1 when the row has more than three
// neighbours), dg: M_ii.  Rows row0 .. row0 + gridDim.x - 1 of P.
template <int NT, int ROWS>
__global__ void __launch_bounds__(NT) k_expm_rows(int n, const int32_t* __restrict__ nbi,
                                                  const double* __restrict__ nbc,
                                                  const int32_t* __restrict__ ovi,
                                                  const double* __restrict__ ovc,
                                                  const double* __restrict__ dg,
                                                  const double* __restrict__ wk, int K,
                                                  const double* __restrict__ C, int row0,
                                                  double* __restrict__ P) {
    __shared__ double buf[2][ROWS * NT];
    const int col = row0 + (int)blockIdx.x;
    uint32_t i01[ROWS];
    int32_t i2[ROWS];
    double a[ROWS][3], d[ROWS], t[ROWS], out[ROWS];
    const double w0 = wk[0];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        const int r = threadIdx.x + q * NT;
        const bool row = r < n;
        i01[q] = row ? ((uint32_t)nbi[3 * r] | ((uint32_t)nbi[3 * r + 1] << 16)) : 0u;
        i2[q] = row ? nbi[3 * r + 2] : 0;
#pragma unroll
        for (int e = 0; e < 3; ++e) a[q][e] = row ? nbc[3 * r + e] : 0.0;
        d[q] = row ? dg[r] : 0.0;
        t[q] = r == col ? 1.0 : 0.0;
        out[q] = w0 * t[q];
        buf[0][r] = t[q];
    }
    __syncthreads();
    int cur = 0;
    for (int k = 1; k <= K; ++k) {
        const double ck = wk[k];
        const double* v = buf[cur];
        double* w = buf[1 - cur];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            const int r = threadIdx.x + q * NT;
            if (r < n) {
                double s = d[q] * t[q];
                s = fma(a[q][0], v[i01[q] & 0xFFFFu], s);
                s = fma(a[q][1], v[i01[q] >> 16], s);
                if (i2[q] >= 0) {
                    s = fma(a[q][2], v[i2[q]], s);
                } else {
                    for (int e = -i2[q] - 1; ovi[e] >= 0; ++e) s = fma(ovc[e], v[ovi[e]], s);
                }
                t[q] = s;
                w[r] = s;
                out[q] = fma(ck, s, out[q]);
            }
        }
        __syncthreads();
        cur = 1 - cur;
    }
    const double inv_cc = 1.0 / C[col];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        const int r = threadIdx.x + q * NT;
        if (r < n) P[(size_t)col * n + r] = out[q] * (C[r] * inv_cc);
    }
}

template <int NT>
static void launch_expm_rows(int rows, int n, const int32_t* nbi, const double* nbc,
                             const int32_t* ovi, const double* ovc, const double* dg,
                             const double* wk, int K, const double* C, double* P, hipStream_t st) {
    const int per = (n + NT - 1) / NT;
    switch (per) {
        case 1: k_expm_rows<NT, 1><<<rows, NT, 0, st>>>(n, nbi, nbc, ovi, ovc, dg, wk, K, C, 0, P); break;
        case 2: k_expm_rows<NT, 2><<<rows, NT, 0, st>>>(n, nbi, nbc, ovi, ovc, dg, wk, K, C, 0, P); break;
        case 3: k_expm_rows<NT, 3><<<rows, NT, 0, st>>>(n, nbi, nbc, ovi, ovc, dg, wk, K, C, 0, P); break;
        case 4: k_expm_rows<NT, 4><<<rows, NT, 0, st>>>(n, nbi, nbc, ovi, ovc, dg, wk, K, C, 0, P); break;
        default: k_expm_rows<NT, 5><<<rows, NT, 0, st>>>(n, nbi, nbc, ovi, ovc, dg, wk, K, C, 0, P); break;
    }
}

// The uniformization operator of one cell on the host: M's neighbour slots (nb, nc, overflow
// lists ovi / ovc) and diagonal dg, and the Poisson weights w[0 .. K]; returns K.
struct SeriesOperator {
    std::vector<int32_t> nb, ovi;
    std::vector<double> nc, ovc, dg, w;
    double sigma = 0.0;
};
static int series_operator(int n, const int32_t* parent, const double* R, const double* C,
                           double dt_ms, SeriesOperator& op) {
    const double dt = dt_ms * 1e-3;  // seconds, electric.py:86
    std::vector<std::vector<std::pair<int, double>>> adj(n);
    std::vector<double> rowsum(n, 0.0);
    for (int c = 0; c < n; ++c) {  // b_ij = dt / (R_edge C_i), electric.py:92-108
        const int q = parent[c];
        if (q < 0) continue;
        NW_REQUIRE(q < n, "parent index %d out of range for cell of %d segments", q, n);
        const double bc = dt / (R[c] * C[c]), bq = dt / (R[c] * C[q]);
        adj[c].push_back({q, bc});
        adj[q].push_back({c, bq});
        rowsum[c] += bc;
        rowsum[q] += bq;
    }
    double smax = 0.0;
    for (double x : rowsum) smax = std::max(smax, x);
    op.sigma = kSeriesMargin * smax;
    const int K = poisson_weights(op.sigma, kSeriesTol, op.w);
    const double inv = op.sigma > 0.0 ? 1.0 / op.sigma : 0.0;
    op.nb.assign(3 * (size_t)n, 0);
    op.nc.assign(3 * (size_t)n, 0.0);
    op.dg.assign(n, 0.0);
    op.ovi.clear();
    op.ovc.clear();
    for (int i = 0; i < n; ++i) {
        const auto& e = adj[i];
        op.dg[i] = 1.0 - rowsum[i] * inv;
        for (int s = 0; s < 3; ++s) op.nb[3 * i + s] = i;
        const int direct = e.size() <= 3 ? (int)e.size() : 2;
        for (int s = 0; s < direct; ++s) {
            op.nb[3 * i + s] = e[s].first;
            op.nc[3 * i + s] = e[s].second * inv;
        }
        if (e.size() > 3) {
            op.nb[3 * i + 2] = -(int32_t)op.ovi.size() - 1;
            for (size_t s = 2; s < e.size(); ++s) {
                op.ovi.push_back(e[s].first);
                op.ovc.push_back(e[s].second * inv);
            }
            op.ovi.push_back(-1);
            op.ovc.push_back(0.0);
        }
    }
    if (op.ovi.empty()) {
        op.ovi.push_back(-1);
        op.ovc.push_back(0.0);
    }
    return K;
}

// Squarings s for a cell of n rows whose B has sigma = kSeriesMargin max|B_ii|: the series of
// B / 2^s costs n workgroups of K(sigma / 2^s) terms, each squaring one FP64 GEMM of 2 n^3 flop;
// s minimises their sum on this device's rates (a 1024-thread workgroup term about
// 300 + 400 ceil(n / 1024) cycles on one of 256 CUs at 2.4 GHz; dgemm about 60 TFLOP/s).
// Squaring non-negative matrices keeps every entry's relative accuracy.
static int series_squarings(int n, double sigma) {
    const int groups = (n + kChebBlock - 1) / kChebBlock;
    const double term_s = (300.0 + 400.0 * groups) / (256.0 * 2.4e9);
    const double gemm_s = 2.0 * (double)n * n * n / 60e12;
    int best = 0;
    double best_cost = 0.0;
    for (int s = 0; s <= 8; ++s) {
        const double sg = std::ldexp(sigma, -s);
        const double cost = n * (sg + 8.5 * std::sqrt(sg) + 10.0) * term_s + s * gemm_s;
        if (s == 0 || cost < best_cost) {
            best = s;
            best_cost = cost;
        }
    }
    return best;
}

// P of one cell (n <= kSeriesMaxRows) by uniformization of B / 2^s and s squarings (rocBLAS
// FP64 GEMMs); *result points at the workspace buffer holding it.
static int expm_cell_series(Workspace& w, int n, const int32_t* parent, const double* R,
                            const double* C, double dt_ms, double** result) {
    NW_REQUIRE(n >= 1 && n <= kSeriesMaxRows, "series build: %d rows", n);
    SeriesOperator op;
    int K = series_operator(n, parent, R, C, dt_ms, op);
    if (K < 0) return K;
    const int squarings = series_squarings(n, op.sigma);
    if (squarings > 0) {
        K = series_operator(n, parent, R, C, std::ldexp(dt_ms, -squarings), op);
        if (K < 0) return K;
    }
    const auto &nb = op.nb, &ovi = op.ovi;
    const auto &nc = op.nc, &ovc = op.ovc, &dg = op.dg, &wts = op.w;
    // device copies in the workspace's series arrays (grown on demand)
    int rc;
    if ((rc = w.series_arrays(n, (int)ovi.size(), K + 1))) return rc;
    NW_HIP(hipMemcpyAsync(w.s_nbi, nb.data(), sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice, w.stream));
    NW_HIP(hipMemcpyAsync(w.s_nbc, nc.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, w.stream));
    NW_HIP(hipMemcpyAsync(w.s_ovi, ovi.data(), sizeof(int32_t) * ovi.size(), hipMemcpyHostToDevice, w.stream));
    NW_HIP(hipMemcpyAsync(w.s_ovc, ovc.data(), sizeof(double) * ovc.size(), hipMemcpyHostToDevice, w.stream));
    NW_HIP(hipMemcpyAsync(w.s_dg, dg.data(), sizeof(double) * n, hipMemcpyHostToDevice, w.stream));
    NW_HIP(hipMemcpyAsync(w.s_c, wts.data(), sizeof(double) * (K + 1), hipMemcpyHostToDevice, w.stream));
    NW_HIP(hipMemcpyAsync(w.s_C, C, sizeof(double) * n, hipMemcpyHostToDevice, w.stream));
    if (n <= 5 * 256)
        launch_expm_rows<256>(n, n, w.s_nbi, w.s_nbc, w.s_ovi, w.s_ovc, w.s_dg, w.s_c, K, w.s_C, w.X[0], w.stream);
    else
        launch_expm_rows<kChebBlock>(n, n, w.s_nbi, w.s_nbc, w.s_ovi, w.s_ovc, w.s_dg, w.s_c, K, w.s_C, w.X[0], w.stream);
    NW_HIP(hipGetLastError());
    double* cur = w.X[0];
    double* nxt = w.X[1];
    for (int q = 0; q < squarings; ++q) {
        if ((rc = gemm(w, n, cur, cur, nxt, 0.0))) return rc;
        std::swap(cur, nxt);
    }
    // the host arrays above must outlive the copies
    NW_HIP(hipStreamSynchronize(w.stream));
    *result = cur;
    return NW_OK;
}

// P of one cell: the series for cells of at most kSeriesMaxRows rows, the dense Taylor path above.
static int exp_cell(Workspace& w, int n, const int32_t* parent, const double* R, const double* C,
                    double dt_ms, double** result) {
    if (n <= kSeriesMaxRows) return expm_cell_series(w, n, parent, R, C, dt_ms, result);
    return expm_cell(w, n, parent, R, C, dt_ms, result);
}

// Preorder slot layout of one cell whose rows have at most three neighbours (adj[i]: the parent,
// an earlier index, first, then the children in increasing order; coefficients already scaled by
// 2 / rho).  Rows go in depth-first preorder, a first child right after its parent, into slots
// s = J t + j (thread t, row group j; the product's contract place of slot s is j * 1024 + t, its
// place in the term vectors cheb_term_place(t, j, J)).  A row reaches the rows in slots s - 1 and
// s + 1 without LDS (own registers, or the neighbouring lane of the same wave) unless s is the
// first slot of a wave (s - 1 is in another wave) or the last (s + 1 is); every other neighbour
// is "far".  Each slot has one LDS read port per term; the row in slot s reads its far
// neighbours through its own port and/or the port of slot s - 1 (same wave).  Rows are placed
// greedily in preorder, a row with one far neighbour taking the previous slot's port when that is
// free (so its own stays free for the next row); a row that finds no port moves on by a hole.  A
// hole goes before a row that is not a first child next to its parent (the slot before it holds
// the last row of another subtree, so no slot neighbour is separated); when the row is such a
// first child, the whole chain of first children it ends moves on by one slot instead (a hole
// before the chain's head), which keeps every parent-child pair of the chain adjacent.  The last
// slot, J * 1024 - 1, stays empty: it holds the zero that unused ports read.
// At least min_rows row groups (one row group takes preorder_layout_pairs' layout instead).
// Returns the slots in use, or -1 when the layout needs more than kChebRows row groups; sets J,
// fills pos (contract place of every natural row) and, per contract place, far (the term-vector
// place its port reads; kChebNoRow set for a slot without a row) and coef (the coefficients of
// the neighbours in slots s - 1 and s + 1, then of the far neighbours read through the slot's
// own port and through the previous slot's port).
static int preorder_layout(int n, const std::vector<std::vector<std::pair<int, double>>>& adj,
                           double scale, int& J_out, std::vector<uint16_t>& pos,
                           std::vector<uint32_t>& far, std::vector<float4>& coef, int min_rows = 2) {
    std::vector<int32_t> place(n, -1), order(n), stack(1, 0);
    int next = 0;
    while (!stack.empty()) {
        const int i = stack.back();
        stack.pop_back();
        place[i] = next;
        order[next++] = i;
        for (int d = (int)adj[i].size() - 1; d >= 0; --d)
            if (adj[i][d].first > i) stack.push_back(adj[i][d].first);
    }
    if (next != n) return -1;
    auto parent_of = [&](int i) { return (!adj[i].empty() && adj[i][0].first < i) ? adj[i][0].first : -1; };
    // preorder position q is the first child of position q - 1
    auto first_child = [&](int q) { return q > 0 && parent_of(order[q]) == order[q - 1]; };
    std::vector<int32_t> slot(n), hole_before(n);
    std::vector<uint8_t> uses(n), port_used;  // uses: 1 = own port, 2 = previous slot's port
    int J = std::max(min_rows, n / kChebBlock + 1), size = -1;
    for (; J <= kChebRows; ++J) {
        const int W = kWave * J, cap = J * kChebBlock - 1;  // slot cap holds the zero
        port_used.assign((size_t)cap + 1, 0);
        std::fill(hole_before.begin(), hole_before.end(), 0);
        int q = 0, s = -1, shifts = 0;
        bool fits = true;
        while (q < n && fits) {
            const int i = order[q], par = parent_of(i);
            const int nch = (int)adj[i].size() - (par >= 0 ? 1 : 0);
            int sn = s + 1 + hole_before[q];
            int shift_from = -1;
            for (;;) {
                if (sn >= cap) {
                    fits = false;
                    break;
                }
                const bool par_near = par >= 0 && slot[place[par]] == sn - 1 && sn % W != 0;
                const int nf = (par >= 0 && !par_near) + (nch >= 1 && (sn + 1) % W == 0) + (nch >= 2);
                const bool prev_free = sn % W != 0 && !port_used[sn - 1];
                if (nf <= 1 || (nf == 2 && prev_free)) {
                    uses[q] = nf == 0 ? 0 : nf == 2 ? 3 : prev_free ? 2 : 1;
                    if (uses[q] & 1) port_used[sn] = 1;
                    if (uses[q] & 2) port_used[sn - 1] = 1;
                    break;
                }
                if (par_near) {  // a first child: moving it alone would part it from its parent
                    shift_from = q;
                    break;
                }
                ++sn;
            }
            if (!fits) break;
            if (shift_from < 0) {
                slot[q] = sn;
                s = sn;
                ++q;
                continue;
            }
            // move the chain of first children that ends at q on by one slot: undo its rows
            int q0 = q - 1;
            while (q0 > 0 && first_child(q0) && slot[q0] == slot[q0 - 1] + 1 && slot[q0] % W != 0) --q0;
            if (++shifts > 8 * (n / kWave + 8)) return -1;
            for (int r = q0; r < q; ++r) {
                if (uses[r] & 1) port_used[slot[r]] = 0;
                if (uses[r] & 2) port_used[slot[r] - 1] = 0;
            }
            ++hole_before[q0];
            q = q0;
            s = q0 > 0 ? slot[q0 - 1] : -1;
        }
        if (fits) {
            size = s + 1;
            break;
        }
    }
    if (size < 0) return -1;
    const int W = kWave * J;
    const uint32_t zero = (uint32_t)cheb_term_place(kChebBlock - 1, J - 1, J);
    J_out = J;
    auto contract_place = [&](int sl) { return (uint32_t)((sl % J) * kChebBlock + sl / J); };
    auto term_place = [&](int sl) { return (uint32_t)cheb_term_place(sl / J, sl % J, J); };
    const size_t base = far.size();
    far.resize(base + (size_t)J * kChebBlock, kChebNoRow | zero);
    coef.resize(base + (size_t)J * kChebBlock, float4{0.0f, 0.0f, 0.0f, 0.0f});
    for (int q = 0; q < n; ++q) {  // rows: the flag off
        const uint32_t pl = contract_place(slot[q]);
        far[base + pl] &= ~kChebNoRow;
    }
    for (int q = 0; q < n; ++q) {
        const int i = order[q];
        const int sl = slot[q];
        const uint32_t pl = contract_place(sl);
        pos[i] = (uint16_t)pl;
        float c4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int nfar = 0;
        for (const auto& e : adj[i]) {
            const int sm = slot[place[e.first]];
            const float a = (float)(scale * e.second);
            if (sm == sl - 1 && sl % W != 0) {
                c4[0] = a;
            } else if (sm == sl + 1 && sm % W != 0) {
                c4[1] = a;
            } else {
                // far: through the previous slot's port first when the row uses it, then its own
                const bool prev = (uses[q] & 2) && nfar == 0;
                const int port = prev ? sl - 1 : sl;
                if (nfar >= 2 || (!prev && !(uses[q] & 1))) return -1;  // placement miscounted
                uint32_t& w = far[base + contract_place(port)];
                w = (w & kChebNoRow) | term_place(sm);
                c4[prev ? 3 : 2] = a;
                ++nfar;
            }
        }
        coef[base + pl] = float4{c4[0], c4[1], c4[2], c4[3]};
    }
    return size;
}

// The one-row-group layout (cells of at most 1,023 slots): the same preorder slots, thread t
// owning slot t, but two LDS reads per row (its far neighbours f and g, the zero place where
// unused) and no port sharing, so a hole is needed only where a row in the last slot of a wave
// would have three far neighbours (parent, first child across the wave, second child).  With
// one row per thread the shared port's DPP would sit on each term's dependency chain, and the
// port layout's holes (about 2 % at one row group, its wave boundaries every 64 slots) would push
// 500-row cells past the 512-thread launch.  far word: f | g << 16 (| kChebNoRow), coef: the
// neighbours in slots s - 1 and s + 1, then f and g.  Returns the slots in use, or -1 when the
// cell needs more than one row group.
static int preorder_layout_pairs(int n, const std::vector<std::vector<std::pair<int, double>>>& adj,
                                 double scale, std::vector<uint16_t>& pos,
                                 std::vector<uint32_t>& far, std::vector<float4>& coef) {
    if (n > kChebBlock - 1) return -1;
    std::vector<int32_t> place(n, -1), order(n), stack(1, 0);
    int next = 0;
    while (!stack.empty()) {
        const int i = stack.back();
        stack.pop_back();
        place[i] = next;
        order[next++] = i;
        for (int d = (int)adj[i].size() - 1; d >= 0; --d)
            if (adj[i][d].first > i) stack.push_back(adj[i][d].first);
    }
    if (next != n) return -1;
    auto parent_of = [&](int i) { return (!adj[i].empty() && adj[i][0].first < i) ? adj[i][0].first : -1; };
    std::vector<int32_t> slot(n);
    int s = -1;
    for (int q = 0; q < n; ++q) {
        int sn = s + 1;
        if (q > 0 && sn % kWave == kWave - 1) {
            const int i = order[q];
            const int par = parent_of(i);
            const int nch = (int)adj[i].size() - (par >= 0 ? 1 : 0);
            const bool left_near = par >= 0 && place[par] == q - 1;
            if ((par >= 0 && !left_near) + (nch >= 1) + (nch >= 2) > 2) ++sn;
        }
        slot[q] = sn;
        s = sn;
    }
    const int size = s + 1;
    if (size > kChebBlock - 1) return -1;
    const uint32_t zero = kChebBlock - 1;  // slot = contract place = term place at one row group
    const size_t base = far.size();
    far.resize(base + kChebBlock, kChebNoRow | zero | (zero << 16));
    coef.resize(base + kChebBlock, float4{0.0f, 0.0f, 0.0f, 0.0f});
    for (int q = 0; q < n; ++q) {
        const int i = order[q];
        const int sl = slot[q], lane = sl % kWave;
        pos[i] = (uint16_t)sl;
        float c4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        uint32_t fp[2] = {zero, zero};
        int nfar = 0;
        for (const auto& e : adj[i]) {
            const int sm = slot[place[e.first]];
            const float a = (float)(scale * e.second);
            if (sm == sl - 1 && lane != 0) {
                c4[0] = a;
            } else if (sm == sl + 1 && lane != kWave - 1) {
                c4[1] = a;
            } else {
                if (nfar >= 2) return -1;
                fp[nfar] = (uint32_t)sm;
                c4[2 + nfar] = a;
                ++nfar;
            }
        }
        far[base + sl] = fp[0] | (fp[1] << 16);
        coef[base + sl] = float4{c4[0], c4[1], c4[2], c4[3]};
    }
    return size;
}

// Chebyshev store: per cell the edge coefficients of B = A dt in the difference form
// (B t)_i = sum_j b_ij (t_j - t_i), b_ij = dt / (R_edge C_i) (electric.py:92-108), scaled by
// 2 / rho, the series coefficients, and the cells ordered by decreasing work (terms x rows).
// Series coefficients below this are dropped (their sum goes to c_0, so constants stay exact).
// The truncated half-step operator is within ~5e-9 of exp(A dt) entrywise; over 1000 steps it
// moves a repeatedly spiking cell's trajectory by a tenth of what FP32 storage of the state moves
// it (1e-8 would move it by as much as FP32 storage, 1e-10 costs 6 % more terms).
static constexpr double kChebTol = 1e-9;

static int build_chebyshev(nw_propagator* p, const int64_t* cell_ptr, const int32_t* parent,
                           const double* R, const double* C, double dt_ms) {
    const int64_t N = p->num_rows;
    const double dt_s = dt_ms * 1e-3;
    std::vector<int32_t> nb(3 * N, 0), ovi;
    std::vector<float> nc(3 * N, 0.0f), ovc;
    p->h_cheb_K.assign(p->num_cells, 0);
    p->h_cheb_off.assign(p->num_cells, 0);
    p->h_rho.assign(p->num_cells, 0.0);
    p->h_cheb_c.clear();
    std::vector<double> coef;
    std::vector<std::vector<std::pair<int, double>>> adj;
    std::vector<double> work;
    // preorder slot layouts (cells whose rows have at most 3 neighbours)
    std::vector<uint16_t> ppos(N, 0);
    std::vector<uint32_t> pfar;
    std::vector<float4> pcoef;
    std::vector<int64_t> pbase(p->num_cells, -1);
    std::vector<int32_t> psize(p->num_cells, 0);
    double terms_rows = 0.0;
    int max_deg = 0;
    int64_t edges = 0;
    for (int64_t cell = 0; cell < p->num_cells; ++cell) {
        const int64_t r0 = cell_ptr[cell];
        const int n = (int)(cell_ptr[cell + 1] - r0);
        adj.assign(n, {});
        std::vector<double> diag(n, 0.0), off(n, 0.0);
        for (int c = 0; c < n; ++c) {
            const int q = parent[r0 + c];
            if (q < 0) continue;
            NW_REQUIRE(q < c, "cell %lld: parent %d of segment %d is not an earlier segment",
                       (long long)cell, q, c);
            const double bc = dt_s / (R[r0 + c] * C[r0 + c]);  // row c, neighbour q
            const double bq = dt_s / (R[r0 + c] * C[r0 + q]);  // row q, neighbour c
            adj[c].push_back({q, bc});
            adj[q].push_back({c, bq});
            diag[c] += bc;
            diag[q] += bq;
            off[c] = -dt_s / (R[r0 + c] * std::sqrt(C[r0 + c] * C[r0 + q]));
            ++edges;
        }
        const double rho = tree_spectral_radius(n, parent + r0, diag.data(), off.data(), work);
        NW_REQUIRE(exp_chebyshev(rho, kChebTol, coef) == 0,
                   "cell %lld: no Chebyshev series for rho = %g", (long long)cell, rho);
        p->h_rho[cell] = rho;
        p->h_cheb_K[cell] = (int32_t)coef.size() - 1;
        p->h_cheb_off[cell] = (int64_t)p->h_cheb_c.size();
        p->h_cheb_c.insert(p->h_cheb_c.end(), coef.begin(), coef.end());
        terms_rows += (double)(coef.size() - 1) * n;
        const double scale = rho > 0.0 ? 2.0 / rho : 0.0;
        int cell_deg = 0;
        for (int i = 0; i < n; ++i) cell_deg = std::max<int>(cell_deg, (int)adj[i].size());
        for (int i = 0; i < n; ++i) ppos[r0 + i] = (uint16_t)i;
        if (cell_deg <= 3 && n > 0 && (int)coef.size() <= kChebCoefCap) {
            std::vector<std::vector<std::pair<int, double>>> sadj(adj);
            for (auto& row : sadj)
                for (auto& e : row) e.second *= scale;
            std::vector<uint16_t> cpos(n);
            const size_t mark = pfar.size();
            int J = 1;
            int size = preorder_layout_pairs(n, sadj, 1.0, cpos, pfar, pcoef);
            if (size < 0) {
                pfar.resize(mark);
                pcoef.resize(mark);
                size = preorder_layout(n, sadj, 1.0, J, cpos, pfar, pcoef);
            }
            if (size > 0) {
                pbase[cell] = (int64_t)mark;
                psize[cell] = size | (J << 16);
                for (int i = 0; i < n; ++i) ppos[r0 + i] = cpos[i];
            } else {
                pfar.resize(mark);
                pcoef.resize(mark);
            }
        }
        for (int i = 0; i < n; ++i) {
            const int64_t g = r0 + i;
            const auto& a = adj[i];
            max_deg = std::max<int>(max_deg, (int)a.size());
            for (int d = 0; d < 3; ++d) nb[3 * g + d] = i;  // unused slots: the row itself, 0
            if (a.size() <= 3) {
                for (size_t d = 0; d < a.size(); ++d) {
                    nb[3 * g + d] = a[d].first;
                    nc[3 * g + d] = (float)(scale * a[d].second);
                }
            } else {
                for (int d = 0; d < 2; ++d) {
                    nb[3 * g + d] = a[d].first;
                    nc[3 * g + d] = (float)(scale * a[d].second);
                }
                nb[3 * g + 2] = -(int32_t)ovi.size() - 1;
                for (size_t d = 2; d < a.size(); ++d) {
                    ovi.push_back(a[d].first);
                    ovc.push_back((float)(scale * a[d].second));
                }
                ovi.push_back(-1);
                ovc.push_back(0.0f);
            }
        }
    }
    if (ovi.empty()) {  // a valid (never read) overflow list
        ovi.push_back(-1);
        ovc.push_back(0.0f);
    }
    p->nnz = 2 * edges;
    p->stored = 0;
    p->max_row_nnz = max_deg;
    p->mean_terms = N ? terms_rows / (double)N : 0.0;
    NW_HIP(hipMalloc(&p->nb_idx, sizeof(int32_t) * 3 * N));
    NW_HIP(hipMalloc(&p->nb_coef, sizeof(float) * 3 * N));
    NW_HIP(hipMalloc(&p->ov_idx, sizeof(int32_t) * ovi.size()));
    NW_HIP(hipMalloc(&p->ov_coef, sizeof(float) * ovc.size()));
    NW_HIP(hipMalloc(&p->cheb_K, sizeof(int32_t) * p->num_cells));
    NW_HIP(hipMalloc(&p->cheb_off, sizeof(int64_t) * p->num_cells));
    NW_HIP(hipMalloc(&p->cheb_c, sizeof(double) * p->h_cheb_c.size()));
    NW_HIP(hipMalloc(&p->cheb_v64, sizeof(double) * N));
    NW_HIP(hipMalloc(&p->step_dev, sizeof(int64_t)));
    NW_HIP(hipMemcpy(p->nb_idx, nb.data(), sizeof(int32_t) * 3 * N, hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->nb_coef, nc.data(), sizeof(float) * 3 * N, hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->ov_idx, ovi.data(), sizeof(int32_t) * ovi.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->ov_coef, ovc.data(), sizeof(float) * ovc.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->cheb_K, p->h_cheb_K.data(), sizeof(int32_t) * p->num_cells,
                     hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->cheb_off, p->h_cheb_off.data(), sizeof(int64_t) * p->num_cells,
                     hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->cheb_c, p->h_cheb_c.data(), sizeof(double) * p->h_cheb_c.size(),
                     hipMemcpyHostToDevice));
    if (pfar.empty()) {  // valid (never read) arrays
        pfar.push_back(kChebNoRow);
        pcoef.push_back(float4{0.0f, 0.0f, 0.0f, 0.0f});
    }
    NW_HIP(hipMalloc(&p->pre_pos, sizeof(uint16_t) * std::max<int64_t>(N, 1)));
    p->pre_places = (int64_t)pfar.size();
    NW_HIP(hipMalloc(&p->pre_far, sizeof(uint32_t) * pfar.size()));
    NW_HIP(hipMalloc(&p->pre_coef, sizeof(float4) * pcoef.size()));
    NW_HIP(hipMalloc(&p->pre_base, sizeof(int64_t) * std::max<int64_t>(p->num_cells, 1)));
    NW_HIP(hipMalloc(&p->pre_size, sizeof(int32_t) * std::max<int64_t>(p->num_cells, 1)));
    NW_HIP(hipMemcpy(p->pre_pos, ppos.data(), sizeof(uint16_t) * N, hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->pre_far, pfar.data(), sizeof(uint32_t) * pfar.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->pre_coef, pcoef.data(), sizeof(float4) * pcoef.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->pre_base, pbase.data(), sizeof(int64_t) * p->num_cells, hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->pre_size, psize.data(), sizeof(int32_t) * p->num_cells, hipMemcpyHostToDevice));
    p->pre_cells = 0;
    p->cheb_maxj = 1;
    p->cheb_small = 1;
    for (int64_t c = 0; c < p->num_cells; ++c) {
        p->pre_cells += pbase[c] >= 0;
        const int64_t n = cell_ptr[c + 1] - cell_ptr[c];
        const int J = pbase[c] >= 0 ? (psize[c] >> 16) : (int)((n + kChebBlock - 1) / kChebBlock);
        p->cheb_maxj = std::max(p->cheb_maxj, J);
        if (pbase[c] < 0 || J != 1 || (psize[c] & 0xFFFF) > kChebSmallBlock) p->cheb_small = 0;
    }
    // launch order: most work first (series terms x rows per thread, then rows)
    std::vector<int32_t> order(p->num_cells);
    std::iota(order.begin(), order.end(), 0);
    auto work_of = [&](int32_t c) {
        const int64_t n = cell_ptr[c + 1] - cell_ptr[c];
        return std::make_pair((int64_t)p->h_cheb_K[c] * ((n + kChebBlock - 1) / kChebBlock), n);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return work_of(a) > work_of(b); });
    NW_HIP(hipMalloc(&p->order, sizeof(int32_t) * p->num_cells));
    NW_HIP(hipMemcpy(p->order, order.data(), sizeof(int32_t) * p->num_cells, hipMemcpyHostToDevice));
    return NW_OK;
}

// Rational store: per cell the rows in level order, the two FP32 edge coefficients of every
// row, and per pole the FP64 pivots of (B - theta I) eliminated from the leaves up.
static int build_rational(nw_propagator* p, const int64_t* cell_ptr, const int32_t* parent,
                          const double* R, const double* C, double dt_ms) {
    const int64_t N = p->num_rows;
    const double dt_s = dt_ms * 1e-3;
    std::vector<uint16_t> ppos(std::max<int64_t>(N, 1), 0);
    std::vector<uint32_t> link(std::max<int64_t>(N, 1), 0u);
    std::vector<float2> bcoef(std::max<int64_t>(N, 1), float2{0.0f, 0.0f});
    std::vector<double2> invd((size_t)kRatPoles * std::max<int64_t>(N, 1), double2{0.0, 0.0});
    std::vector<uint16_t> levs;
    std::vector<int64_t> lev_off(p->num_cells, 0), pbase(p->num_cells, 0);
    std::vector<int32_t> depth(p->num_cells, 0), psize(p->num_cells, 0);
    std::vector<int32_t> order_lv, first, count, nat_of;
    std::vector<std::vector<int32_t>> kids;
    double lev_rows = 0.0;
    int max_levels = 0, max_deg = 0;
    int64_t edges = 0;
    for (int64_t cell = 0; cell < p->num_cells; ++cell) {
        const int64_t r0 = cell_ptr[cell];
        const int n = (int)(cell_ptr[cell + 1] - r0);
        kids.assign(n, {});
        for (int c = 1; c < n; ++c) {
            const int q = parent[r0 + c];
            NW_REQUIRE(q >= 0 && q < c, "cell %lld: segment %d needs an earlier parent (one root, "
                       "segment 0)", (long long)cell, c);
            kids[q].push_back(c);
        }
        NW_REQUIRE(parent[r0] < 0, "cell %lld: segment 0 must be the root", (long long)cell);
        // level order: breadth first, a parent's children adjacent in segment order
        nat_of.assign(1, 0);
        std::vector<int32_t> place(n, 0);
        lev_off[cell] = (int64_t)levs.size();
        levs.push_back(0);
        int lo = 0, hi = 1;
        while (lo < hi) {
            for (int k = lo; k < hi; ++k)
                for (int c : kids[nat_of[k]]) {
                    place[c] = (int)nat_of.size();
                    nat_of.push_back(c);
                }
            levs.push_back((uint16_t)hi);
            lo = hi;
            hi = (int)nat_of.size();
        }
        NW_REQUIRE((int)nat_of.size() == n, "cell %lld: the parents do not form one tree",
                   (long long)cell);
        const int D = (int)(levs.size() - lev_off[cell]) - 1;
        NW_REQUIRE(D <= 4095, "cell %lld: %d levels (the rational store holds 4095)",
                   (long long)cell, D);
        depth[cell] = D;
        max_levels = std::max(max_levels, D);
        lev_rows += (double)D * n;
        std::vector<double> bup(n, 0.0), bdn(n, 0.0), diag(n, 0.0);
        for (int k = 0; k < n; ++k) {
            const int i = nat_of[k];
            ppos[r0 + i] = (uint16_t)k;
            const int cnt = (int)kids[i].size();
            NW_REQUIRE(cnt <= 63, "cell %lld: segment %d has %d children (the rational store holds "
                       "63)", (long long)cell, i, cnt);
            max_deg = std::max(max_deg, cnt + (i > 0 ? 1 : 0));
            const int fc = cnt ? place[kids[i][0]] : 0;
            const int par = i > 0 ? place[parent[r0 + i]] : 0;
            link[r0 + k] = (uint32_t)par | ((uint32_t)fc << 13) | ((uint32_t)cnt << 26);
            if (i > 0) {
                const int q = parent[r0 + i];
                // B[parent][row] and B[row][parent] (electric.py:92-108, times the half step)
                const float up = (float)(dt_s / (R[r0 + i] * C[r0 + q]));
                const float dn = (float)(dt_s / (R[r0 + i] * C[r0 + i]));
                bcoef[r0 + k] = float2{up, dn};
                bup[k] = up;
                bdn[k] = dn;
                ++edges;
            }
        }
        // zero row sums from the FP32 coefficients: B[i][i] = -(B[i][parent] + sum B[i][child])
        for (int k = 0; k < n; ++k) {
            const uint32_t L = link[r0 + k];
            const int fc = (int)((L >> 13) & 0x1FFFu), cnt = (int)(L >> 26);
            double d = -bdn[k];
            for (int c = fc; c < fc + cnt; ++c) d -= bup[c];
            diag[k] = d;
        }
        for (int pj = 0; pj < kRatPoles; ++pj) {
            const std::complex<double> th(kRatHost[1 + 4 * pj], kRatHost[2 + 4 * pj]);
            std::vector<std::complex<double>> iv(n);
            for (int k = n - 1; k >= 0; --k) {
                const uint32_t L = link[r0 + k];
                const int fc = (int)((L >> 13) & 0x1FFFu), cnt = (int)(L >> 26);
                std::complex<double> d = diag[k] - th;
                for (int c = fc; c < fc + cnt; ++c) d -= bup[c] * bdn[c] * iv[c];
                iv[k] = 1.0 / d;
                invd[(size_t)pj * N + r0 + k] = double2{iv[k].real(), iv[k].imag()};
            }
        }
        const int J = (n + kChebBlock - 1) / kChebBlock;
        psize[cell] = n | (J << 16);
    }
    p->nnz = 2 * edges;
    p->stored = 0;
    p->max_row_nnz = max_deg;
    p->mean_terms = 0.0;
    p->rat_max_levels = max_levels;
    p->rat_mean_levels = N ? lev_rows / (double)N : 0.0;
    // the Chebyshev arrays the kernels address but never read for this store
    p->h_cheb_K.assign(p->num_cells, 0);
    p->h_cheb_off.assign(p->num_cells, 0);
    p->h_cheb_c.assign(1, 1.0);
    p->h_rho.assign(p->num_cells, 0.0);
    NW_HIP(hipMalloc(&p->nb_idx, sizeof(int32_t) * 3));
    NW_HIP(hipMalloc(&p->nb_coef, sizeof(float) * 3));
    NW_HIP(hipMalloc(&p->ov_idx, sizeof(int32_t)));
    NW_HIP(hipMalloc(&p->ov_coef, sizeof(float)));
    NW_HIP(hipMalloc(&p->cheb_K, sizeof(int32_t) * p->num_cells));
    NW_HIP(hipMalloc(&p->cheb_off, sizeof(int64_t) * p->num_cells));
    NW_HIP(hipMalloc(&p->cheb_c, sizeof(double)));
    NW_HIP(hipMalloc(&p->pre_far, sizeof(uint32_t)));
    NW_HIP(hipMalloc(&p->pre_coef, sizeof(float4)));
    NW_HIP(hipMemcpy(p->cheb_K, p->h_cheb_K.data(), sizeof(int32_t) * p->num_cells,
                     hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->cheb_off, p->h_cheb_off.data(), sizeof(int64_t) * p->num_cells,
                     hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->cheb_c, p->h_cheb_c.data(), sizeof(double), hipMemcpyHostToDevice));
    NW_HIP(hipMalloc(&p->cheb_v64, sizeof(double) * std::max<int64_t>(N, 1)));
    NW_HIP(hipMalloc(&p->step_dev, sizeof(int64_t)));
    NW_HIP(hipMalloc(&p->pre_pos, sizeof(uint16_t) * ppos.size()));
    NW_HIP(hipMalloc(&p->pre_base, sizeof(int64_t) * p->num_cells));
    NW_HIP(hipMalloc(&p->pre_size, sizeof(int32_t) * p->num_cells));
    NW_HIP(hipMalloc(&p->rat_link, sizeof(uint32_t) * link.size()));
    NW_HIP(hipMalloc(&p->rat_b, sizeof(float2) * bcoef.size()));
    NW_HIP(hipMalloc(&p->rat_invd, sizeof(double2) * invd.size()));
    NW_HIP(hipMalloc(&p->rat_lev, sizeof(uint16_t) * levs.size()));
    NW_HIP(hipMalloc(&p->rat_lev_off, sizeof(int64_t) * p->num_cells));
    NW_HIP(hipMalloc(&p->rat_depth, sizeof(int32_t) * p->num_cells));
    NW_HIP(hipMemcpy(p->pre_pos, ppos.data(), sizeof(uint16_t) * ppos.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->pre_base, pbase.data(), sizeof(int64_t) * p->num_cells, hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->pre_size, psize.data(), sizeof(int32_t) * p->num_cells, hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->rat_link, link.data(), sizeof(uint32_t) * link.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->rat_b, bcoef.data(), sizeof(float2) * bcoef.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->rat_invd, invd.data(), sizeof(double2) * invd.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->rat_lev, levs.data(), sizeof(uint16_t) * levs.size(), hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->rat_lev_off, lev_off.data(), sizeof(int64_t) * p->num_cells,
                     hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->rat_depth, depth.data(), sizeof(int32_t) * p->num_cells, hipMemcpyHostToDevice));
    p->pre_cells = p->num_cells;
    p->pre_places = N;
    p->cheb_maxj = 1;
    p->cheb_small = 1;
    for (int64_t c = 0; c < p->num_cells; ++c) {
        const int n = psize[c] & 0xFFFF, J = psize[c] >> 16;
        p->cheb_maxj = std::max(p->cheb_maxj, J);
        if (n > kChebSmallBlock) p->cheb_small = 0;
    }
    // launch order: most work first (levels x rows per thread, then rows)
    std::vector<int32_t> order(p->num_cells);
    std::iota(order.begin(), order.end(), 0);
    auto work_of = [&](int32_t c) {
        const int64_t n = cell_ptr[c + 1] - cell_ptr[c];
        return std::make_pair((int64_t)depth[c] * ((n + kChebBlock - 1) / kChebBlock) + n / 64, n);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return work_of(a) > work_of(b); });
    NW_HIP(hipMalloc(&p->order, sizeof(int32_t) * p->num_cells));
    NW_HIP(hipMemcpy(p->order, order.data(), sizeof(int32_t) * p->num_cells, hipMemcpyHostToDevice));
    return NW_OK;
}

static void free_prop(nw_propagator* p) {
    if (!p) return;
    for (auto& g : p->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    p->graphs.clear();
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    if (p->step_dev) (void)hipFree(p->step_dev);
    if (p->nb_idx) (void)hipFree(p->nb_idx);
    if (p->nb_coef) (void)hipFree(p->nb_coef);
    if (p->ov_idx) (void)hipFree(p->ov_idx);
    if (p->ov_coef) (void)hipFree(p->ov_coef);
    if (p->cheb_K) (void)hipFree(p->cheb_K);
    if (p->cheb_off) (void)hipFree(p->cheb_off);
    if (p->cheb_c) (void)hipFree(p->cheb_c);
    if (p->cheb_v64) (void)hipFree(p->cheb_v64);
    if (p->pre_pos) (void)hipFree(p->pre_pos);
    if (p->pre_far) (void)hipFree(p->pre_far);
    if (p->pre_coef) (void)hipFree(p->pre_coef);
    if (p->pre_base) (void)hipFree(p->pre_base);
    if (p->pre_size) (void)hipFree(p->pre_size);
    if (p->rat_link) (void)hipFree(p->rat_link);
    if (p->rat_b) (void)hipFree(p->rat_b);
    if (p->rat_invd) (void)hipFree(p->rat_invd);
    if (p->rat_lev) (void)hipFree(p->rat_lev);
    if (p->rat_lev_off) (void)hipFree(p->rat_lev_off);
    if (p->rat_depth) (void)hipFree(p->rat_depth);
    if (p->cell_row0) (void)hipFree(p->cell_row0);
    if (p->cell_nz0) (void)hipFree(p->cell_nz0);
    if (p->row_off) (void)hipFree(p->row_off);
    if (p->vals) (void)hipFree(p->vals);
    if (p->cols) (void)hipFree(p->cols);
    if (p->rowsum) (void)hipFree(p->rowsum);
    if (p->wgt) (void)hipFree(p->wgt);
    if (p->order) (void)hipFree(p->order);
    delete p;
}

static int build(int device, int64_t num_cells, const int64_t* cell_ptr, const int32_t* parent,
                 const double* R, const double* C, double dt_ms, double prune, int flags,
                 nw_propagator* p) {
    const bool sym = (flags & NW_PROPAGATOR_SYMMETRIC) != 0;
    const bool rat = (flags & NW_PROPAGATOR_RATIONAL) != 0;
    const bool cheb = (flags & NW_PROPAGATOR_CHEBYSHEV) != 0 || rat;
    NW_REQUIRE((int)sym + (int)((flags & NW_PROPAGATOR_CHEBYSHEV) != 0) + (int)rat <= 1,
               "the symmetric, Chebyshev and rational stores are exclusive");
    p->symmetric = sym ? 1 : 0;
    p->chebyshev = cheb ? 1 : 0;
    p->rational = rat ? 1 : 0;
    auto t0 = std::chrono::steady_clock::now();
    NW_REQUIRE(num_cells > 0, "num_cells must be positive");
    NW_REQUIRE(cell_ptr[0] == 0, "cell_ptr[0] must be 0");
    NW_REQUIRE(dt_ms > 0.0 && prune >= 0.0, "time step must be positive and prune non-negative");
    int nmax = 0;
    const int64_t limit = cheb ? (int64_t)kChebBlock * kChebRows : 65535;
    for (int64_t c = 0; c < num_cells; ++c) {
        int64_t n = cell_ptr[c + 1] - cell_ptr[c];
        NW_REQUIRE(n >= 1 && n <= limit, "cell %lld has %lld segments (1..%lld supported%s)",
                   (long long)c, (long long)n, (long long)limit,
                   rat ? " by the rational store" : cheb ? " by the Chebyshev store" : "");
        nmax = std::max<int>(nmax, (int)n);
    }
    NW_HIP(hipSetDevice(device));
    p->device = device;
    p->num_cells = num_cells;
    p->num_rows = cell_ptr[num_cells];
    p->max_cell_rows = nmax;
    p->h_cell_row0.assign(cell_ptr, cell_ptr + num_cells + 1);
    p->h_cell_nz0.assign(num_cells, 0);
    if (cheb) {
        NW_HIP(hipMalloc(&p->cell_row0, sizeof(int64_t) * (num_cells + 1)));
        NW_HIP(hipMemcpy(p->cell_row0, cell_ptr, sizeof(int64_t) * (num_cells + 1),
                         hipMemcpyHostToDevice));
        int rc = rat ? build_rational(p, cell_ptr, parent, R, C, dt_ms)
                     : build_chebyshev(p, cell_ptr, parent, R, C, dt_ms);
        p->build_seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    }
    NW_HIP(hipMalloc(&p->row_off, sizeof(uint32_t) * (p->num_rows + num_cells)));
    NW_HIP(hipMalloc(&p->rowsum, sizeof(double) * p->num_rows));
    if (sym) NW_HIP(hipMalloc(&p->wgt, sizeof(double) * p->num_rows));
    // capacity guess (about 540 stored entries per row at prune 1e-9, half that symmetric);
    // grown on demand
    int64_t guess = 0;
    for (int64_t c = 0; c < num_cells; ++c) {
        int64_t n = cell_ptr[c + 1] - cell_ptr[c];
        guess += n * std::min<int64_t>(n, sym ? 320 : 640);
    }
    int rc;
    if ((rc = grow(p, std::max<int64_t>(guess, 1)))) return rc;

    Workspace w;
    if ((rc = w.init(nmax, nmax > kSeriesMaxRows))) return rc;
    std::vector<uint32_t> off;
    std::vector<int64_t> cell_nnz(num_cells);
    for (int64_t c = 0; c < num_cells; ++c) {
        int64_t r0 = cell_ptr[c];
        int n = (int)(cell_ptr[c + 1] - r0);
        double* P = nullptr;
        if ((rc = exp_cell(w, n, parent + r0, R + r0, C + r0, dt_ms, &P))) return rc;
        if (sym) {  // S = W P W^-1, w = sqrt(C / C_first) (any per-cell scale cancels)
            std::vector<double> wv(n);
            for (int i = 0; i < n; ++i) wv[i] = std::sqrt(C[r0 + i] / C[r0]);
            NW_HIP(hipMemcpyAsync(w.d_w, wv.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                  w.stream));
            double* Sd = (P == w.X[0]) ? w.X[1] : w.X[0];
            dim3 blk(32, 8), grd((n + 31) / 32, (n + 31) / 32);
            k_symmetrize<<<grd, blk, 0, w.stream>>>(P, w.d_w, n, Sd);
            NW_HIP(hipGetLastError());
            P = Sd;
        }
        int32_t max_row = 0;
        int64_t kept = 0;
        if ((rc = count_rows(w, P, n, prune, kRowAlign, off, max_row, kept, sym))) return rc;
        p->max_row_nnz = std::max(p->max_row_nnz, max_row);
        int64_t cnt = off[n];
        if ((rc = grow(p, p->stored + cnt))) return rc;
        uint32_t* dev_off = p->row_off + r0 + c;
        NW_HIP(hipMemcpyAsync(dev_off, off.data(), (n + 1) * sizeof(uint32_t),
                              hipMemcpyHostToDevice, w.stream));
        if (sym)
            k_fill_sym<<<(n + 3) / 4, 4 * kWave, 0, w.stream>>>(
                P, w.d_w, n, prune, dev_off, p->vals + p->stored, p->cols + p->stored,
                p->rowsum + r0, p->wgt + r0);
        else
            k_fill<true><<<(n + 3) / 4, 4 * kWave, 0, w.stream>>>(
                P, n, prune, dev_off, p->vals + p->stored, p->cols + p->stored, p->rowsum + r0,
                nullptr, nullptr);
        NW_HIP(hipGetLastError());
        p->h_cell_nz0[c] = p->stored;
        cell_nnz[c] = cnt;
        p->stored += cnt;
        p->nnz += kept;
    }
    NW_HIP(hipStreamSynchronize(w.stream));
    NW_HIP(hipMalloc(&p->cell_row0, sizeof(int64_t) * (num_cells + 1)));
    NW_HIP(hipMalloc(&p->cell_nz0, sizeof(int64_t) * num_cells));
    NW_HIP(hipMemcpy(p->cell_row0, cell_ptr, sizeof(int64_t) * (num_cells + 1),
                     hipMemcpyHostToDevice));
    NW_HIP(hipMemcpy(p->cell_nz0, p->h_cell_nz0.data(), sizeof(int64_t) * num_cells,
                     hipMemcpyHostToDevice));
    // launch order: largest stored blocks first (longest-processing-time scheduling)
    std::vector<int32_t> order(num_cells);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return cell_nnz[a] > cell_nnz[b]; });
    NW_HIP(hipMalloc(&p->order, sizeof(int32_t) * num_cells));
    NW_HIP(hipMemcpy(p->order, order.data(), sizeof(int32_t) * num_cells, hipMemcpyHostToDevice));
    p->build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return NW_OK;
}

// ----------------------------------------------------------------------------------------------
// Half step
// ----------------------------------------------------------------------------------------------
struct HalfArgs {
    const int64_t* cell_row0;
    const int64_t* cell_nz0;
    const uint32_t* row_off;
    const float* vals;
    const uint16_t* cols;
    const double* rowsum;
    const int32_t* order;
    int32_t nmax;
    float half_dt_ms;
    float* v;
    const float* sum_g;
    const float* drive_v;
    const float* sum_i;
    const float* inv_c;
    float* integral_v;
    const double* wgt;  // symmetric store: w of every segment
};

// FP32 storage, FP64 arithmetic: x is formed in FP64 and staged in LDS as an FP32 offset from the
// cell's reference voltage (x of its first segment); the row products accumulate in FP64
// (spmv_rows).
__device__ __forceinline__ double half_x_of(const HalfArgs& a, int64_t gi, double v, double d,
                                            double sum_i, double sum_g, double invc, double dt_ms,
                                            double dt_s, bool write_integral) {
    const double dv = dt_ms * sum_i * invc;
    const double expo = -dt_s * sum_g * invc;
    const double alpha = exp(expo);
    const double diff = d - v;
    if (write_integral && a.integral_v)
        a.integral_v[gi] = (float)(dt_s * d - expo * diff * alpha + 0.5 * dt_s * dv);
    return d - diff * alpha + dv;
}

__device__ __forceinline__ double half_x(const HalfArgs& a, int64_t gi, double dt_ms, double dt_s,
                                         bool write_integral) {
    return half_x_of(a, gi, a.v[gi], a.drive_v[gi], a.sum_i[gi], a.sum_g[gi], a.inv_c[gi], dt_ms,
                     dt_s, write_integral);
}

// SYM = false: x - vref and the product in LDS as FP32 (spmv_rows).  SYM = true (symmetric store):
// y = W (x - vref) in LDS as FP32, z = S y as FP64 (spmv_sym_rows), v = vref * rowsum + z / w.
template <bool SYM>
__global__ void __launch_bounds__(256) k_half_step(HalfArgs a) {
    extern __shared__ double lds64[];
    __shared__ double s_vref;
    float* xs = SYM ? reinterpret_cast<float*>(lds64 + a.nmax) : reinterpret_cast<float*>(lds64);
    double* zs = SYM ? lds64 : nullptr;
    const int cell = a.order[blockIdx.x];
    const int64_t row0 = a.cell_row0[cell];
    const int n = (int)(a.cell_row0[cell + 1] - row0);
    const double dt_ms = a.half_dt_ms;
    const double dt_s = dt_ms * 1e-3;
    const double* wg = SYM ? a.wgt + row0 : nullptr;
    if (threadIdx.x == 0) s_vref = half_x(a, row0, dt_ms, dt_s, false);
    __syncthreads();
    const double vref = s_vref;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double x = half_x(a, row0 + i, dt_ms, dt_s, true) - vref;
        if (SYM) {
            xs[i] = (float)(x * wg[i]);
            zs[i] = 0.0;
        } else {
            xs[i] = (float)x;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nwaves = blockDim.x / kWave;
    const uint32_t* roff = a.row_off + row0 + cell;
    const float* vals = a.vals + a.cell_nz0[cell];
    const uint16_t* cols = a.cols + a.cell_nz0[cell];
    const double* rs = a.rowsum + row0;
    float* vout = xs + a.nmax;
    if (SYM)
        spmv_sym_rows(roff, vals, cols, xs, zs, n, wave, nwaves, lane);
    else
        spmv_rows(roff, vals, cols, rs, xs, vout, vref, n, wave, nwaves, lane);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = SYM ? fma(vref, rs[i], zs[i] / wg[i]) : vref + (double)vout[i];
        a.v[row0 + i] = (float)v;
    }
}

// Chebyshev store: x in LDS as FP64, v = sum_k c_k T_k(A~) x (cheb_rows).
template <int MAXJ, int NT = kChebBlock, int KIND = kProdAny>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MAXJ == 1 ? 8 : 4)))
k_half_step_cheb(HalfArgs a, ChebStore cs) {
    __shared__ __align__(16) double b0[prod_lds(MAXJ, KIND)];
    const int cell = a.order[blockIdx.x];
    const int64_t row0 = a.cell_row0[cell];
    const int n = (int)(a.cell_row0[cell + 1] - row0);
    const double dt_ms = a.half_dt_ms;
    const double dt_s = dt_ms * 1e-3;
    const int64_t pre_base = cs.pre_base[cell];
    const bool pre = pre_base >= 0;
    const uint16_t* pos = cs.pre_pos + row0;
    // the thread's rows (i = threadIdx.x + j * 1024, j < MAXJ) in groups of at most three whose
    // loads are issued together, then the arithmetic
    constexpr int G = MAXJ < 3 ? MAXJ : 3;
#pragma unroll
    for (int j0 = 0; j0 < MAXJ; j0 += G) {
        float fv[G], fd[G], fi[G], fg[G], fc[G];
        int place[G];
#pragma unroll
        for (int jj = 0; jj < G; ++jj) {
            const int i = threadIdx.x + (j0 + jj) * kChebBlock;
            if (j0 + jj < MAXJ && i < n) {
                const int64_t gi = row0 + i;
                fv[jj] = a.v[gi];
                fd[jj] = a.drive_v[gi];
                fi[jj] = a.sum_i[gi];
                fg[jj] = a.sum_g[gi];
                fc[jj] = a.inv_c[gi];
                place[jj] = pre ? (int)pos[i] : i;
            }
        }
#pragma unroll
        for (int jj = 0; jj < G; ++jj) {
            const int i = threadIdx.x + (j0 + jj) * kChebBlock;
            if (j0 + jj >= MAXJ || i >= n) continue;
            b0[place[jj]] = half_x_of(a, row0 + i, fv[jj], fd[jj], fi[jj], fg[jj], fc[jj], dt_ms,
                                      dt_s, true);
        }
    }
    __syncthreads();
    double out[kChebRows];
    const int places = cheb_cell<MAXJ, NT, KIND>(cs, row0, cell, n, pre_base, b0, out);
#pragma unroll
    for (int j = 0; j < kChebRows; ++j) {
        const int r = threadIdx.x + j * kChebBlock;
        if (r < places) b0[r] = out[j];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NT) a.v[row0 + i] = (float)b0[pre ? pos[i] : i];
}

}  // namespace nw

using namespace nw;

extern "C" {

const char* nw_last_error(void) { return nw::g_err; }

const char* nw_version(void) { return "neuwon_amd 0.1 (gfx950)"; }

int nw_device_info(int device, char* name_out, int name_len, int* num_cus, int* lds_bytes) {
    hipDeviceProp_t prop;
    NW_HIP(hipGetDeviceProperties(&prop, device));
    if (name_out && name_len > 0) {
        snprintf(name_out, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (lds_bytes) *lds_bytes = (int)prop.sharedMemPerBlock;
    return NW_OK;
}

uint32_t nw_stim_hash(uint32_t seed, uint32_t cell, uint32_t tick) {
    return nw::stim_hash(seed, cell, tick);
}

int nw_propagator_build(int device, int64_t num_cells, const int64_t* cell_ptr_host,
                        const int32_t* parent_local_host, const double* R_host,
                        const double* C_host, double time_step_ms, double prune,
                        nw_propagator** out) {
    return nw_propagator_build_ex(device, num_cells, cell_ptr_host, parent_local_host, R_host,
                                  C_host, time_step_ms, prune, 0, out);
}

int nw_propagator_build_ex(int device, int64_t num_cells, const int64_t* cell_ptr_host,
                           const int32_t* parent_local_host, const double* R_host,
                           const double* C_host, double time_step_ms, double prune, int32_t flags,
                           nw_propagator** out) {
    NW_REQUIRE(out && cell_ptr_host && parent_local_host && R_host && C_host,
               "nw_propagator_build: null argument");
    NW_REQUIRE((flags & ~(NW_PROPAGATOR_SYMMETRIC | NW_PROPAGATOR_CHEBYSHEV | NW_PROPAGATOR_RATIONAL)) == 0,
               "unknown propagator flags %d", flags);
    *out = nullptr;
    auto* p = new nw_propagator();
    int rc = build(device, num_cells, cell_ptr_host, parent_local_host, R_host, C_host,
                   time_step_ms, prune, flags, p);
    if (rc) {
        free_prop(p);
        return rc;
    }
    *out = p;
    return NW_OK;
}

int nw_propagator_stats_get(const nw_propagator* p, nw_propagator_stats* s) {
    NW_REQUIRE(p && s, "nw_propagator_stats_get: null argument");
    s->num_cells = p->num_cells;
    s->num_rows = p->num_rows;
    s->nnz = p->nnz;
    s->stored = p->stored;
    s->max_cell_rows = p->max_cell_rows;
    s->max_row_nnz = p->max_row_nnz;
    s->build_seconds = p->build_seconds;
    s->symmetric = p->symmetric;
    s->chebyshev = p->chebyshev;
    s->mean_terms = p->mean_terms;
    s->workgroup_threads = p->chebyshev ? (p->cheb_small ? kChebSmallBlock : kChebBlock) : 0;
    s->row_groups = p->chebyshev ? (p->cheb_maxj == 1 ? 1 : kChebRows) : 0;
    s->rational = p->rational;
    s->max_levels = p->rat_max_levels;
    s->mean_levels = p->rat_mean_levels;
    s->device_bytes = p->vals_capacity * (int64_t)(sizeof(float) + sizeof(uint16_t)) +
                      (p->num_rows + p->num_cells) * (int64_t)sizeof(uint32_t) +
                      p->num_rows * (int64_t)sizeof(double) * (p->symmetric ? 2 : 1) +
                      p->num_cells * (int64_t)(2 * sizeof(int64_t) + sizeof(int32_t));
    if (p->chebyshev)  // natural-order neighbours, preorder slots, series, FP64 voltages
        s->device_bytes = p->num_rows * (int64_t)(3 * sizeof(int32_t) + 3 * sizeof(float) +
                                                  sizeof(uint16_t) + sizeof(double)) +
                          p->pre_places * (int64_t)(sizeof(uint32_t) + sizeof(float4)) +
                          (int64_t)p->h_cheb_c.size() * (int64_t)sizeof(double) +
                          p->num_cells * (int64_t)(3 * sizeof(int64_t) + 3 * sizeof(int32_t));
    if (p->rational)  // level places, links, edge pairs, pivots, FP64 voltages, level starts
        s->device_bytes = p->num_rows * (int64_t)(sizeof(uint16_t) + sizeof(uint32_t) + sizeof(float2) +
                                                  kRatPoles * sizeof(double2) + sizeof(double)) +
                          (p->num_rows + 2 * p->num_cells) * (int64_t)sizeof(uint16_t) +
                          p->num_cells * (int64_t)(4 * sizeof(int64_t) + 4 * sizeof(int32_t));
    return NW_OK;
}

int nw_propagator_cell_chebyshev(const nw_propagator* p, int64_t cell, int32_t* K_out,
                                 double* rho_out, double* coef_host) {
    NW_REQUIRE(p && K_out && rho_out, "nw_propagator_cell_chebyshev: null argument");
    NW_REQUIRE(p->chebyshev && !p->rational, "the propagator is not in the Chebyshev store");
    NW_REQUIRE(cell >= 0 && cell < p->num_cells, "cell index out of range");
    *K_out = p->h_cheb_K[cell];
    *rho_out = p->h_rho[cell];
    if (coef_host)
        std::memcpy(coef_host, p->h_cheb_c.data() + p->h_cheb_off[cell],
                    sizeof(double) * (p->h_cheb_K[cell] + 1));
    return NW_OK;
}

int nw_chebyshev_exp_coefficients(double rho, double tol, double* coef_host, int32_t capacity) {
    NW_REQUIRE(rho >= 0.0 && tol > 0.0 && capacity >= 0, "bad argument");
    std::vector<double> c;
    NW_REQUIRE(exp_chebyshev(rho, tol, c) == 0, "no series for rho = %g", rho);
    if (coef_host)
        std::memcpy(coef_host, c.data(), sizeof(double) * std::min<size_t>(c.size(), capacity));
    return (int)c.size();
}

int nw_rational_coefficients(double* out, int32_t capacity) {
    NW_REQUIRE(out && capacity >= 1 + 4 * kRatPoles, "nw_rational_coefficients: room for %d values",
               1 + 4 * kRatPoles);
    std::memcpy(out, kRatHost, sizeof(kRatHost));
    return 1 + 4 * kRatPoles;
}

int nw_propagator_destroy(nw_propagator* p) {
    free_prop(p);
    return NW_OK;
}

int nw_propagator_cell_f64(int device, int32_t n, const int32_t* parent_local_host,
                           const double* R_host, const double* C_host, double time_step_ms,
                           double prune, int64_t* nnz_out, int64_t* row_ptr_host,
                           int32_t* col_host, double* val_host) {
    NW_REQUIRE(n >= 1 && n <= 65535, "cell size %d out of range", n);
    NW_REQUIRE(parent_local_host && R_host && C_host && nnz_out, "null argument");
    NW_HIP(hipSetDevice(device));
    Workspace w;
    int rc;
    if ((rc = w.init(n, n > kSeriesMaxRows))) return rc;
    double* P = nullptr;
    if ((rc = exp_cell(w, n, parent_local_host, R_host, C_host, time_step_ms, &P))) return rc;
    std::vector<uint32_t> off;
    int32_t max_row;
    int64_t kept;
    if ((rc = count_rows(w, P, n, prune, 1, off, max_row, kept))) return rc;
    *nnz_out = off[n];
    if (!val_host) return NW_OK;
    NW_REQUIRE(row_ptr_host && col_host, "null output array");
    double* dv = nullptr;
    int32_t* dc = nullptr;
    NW_HIP(hipMalloc(&dv, sizeof(double) * std::max<int64_t>(off[n], 1)));
    NW_HIP(hipMalloc(&dc, sizeof(int32_t) * std::max<int64_t>(off[n], 1)));
    NW_HIP(hipMemcpyAsync(w.d_off, off.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                          w.stream));
    k_fill<false><<<(n + 3) / 4, 4 * kWave, 0, w.stream>>>(P, n, prune, w.d_off, nullptr, nullptr,
                                                           nullptr, dv, dc);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(val_host, dv, sizeof(double) * off[n],
                                            hipMemcpyDeviceToHost, w.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(col_host, dc, sizeof(int32_t) * off[n],
                                            hipMemcpyDeviceToHost, w.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w.stream);
    (void)hipFree(dv);
    (void)hipFree(dc);
    if (e != hipSuccess) {
        set_error("nw_propagator_cell_f64: %s", hipGetErrorString(e));
        return NW_ERR_HIP;
    }
    for (int i = 0; i <= n; ++i) row_ptr_host[i] = off[i];
    return NW_OK;
}

int nw_propagator_cell_weights(const nw_propagator* p, int64_t cell, double* w_host) {
    NW_REQUIRE(p && w_host && cell >= 0 && cell < p->num_cells, "bad argument");
    NW_REQUIRE(p->symmetric, "the propagator is not in the symmetric format");
    int64_t r0 = p->h_cell_row0[cell];
    int n = (int)(p->h_cell_row0[cell + 1] - r0);
    NW_HIP(hipMemcpy(w_host, p->wgt + r0, n * sizeof(double), hipMemcpyDeviceToHost));
    return NW_OK;
}

int nw_propagator_cell_export(const nw_propagator* p, int64_t cell, uint32_t* row_off_host,
                              uint16_t* col_host, float* val_host) {
    NW_REQUIRE(p && cell >= 0 && cell < p->num_cells, "cell index out of range");
    NW_REQUIRE(!p->chebyshev, "the Chebyshev store holds no matrix to export");
    NW_HIP(hipSetDevice(p->device));
    int64_t r0 = p->h_cell_row0[cell];
    int n = (int)(p->h_cell_row0[cell + 1] - r0);
    std::vector<uint32_t> off(n + 1);
    NW_HIP(hipMemcpy(off.data(), p->row_off + r0 + cell, (n + 1) * sizeof(uint32_t),
                     hipMemcpyDeviceToHost));
    if (row_off_host) std::memcpy(row_off_host, off.data(), (n + 1) * sizeof(uint32_t));
    int64_t z0 = p->h_cell_nz0[cell];
    if (col_host)
        NW_HIP(hipMemcpy(col_host, p->cols + z0, off[n] * sizeof(uint16_t), hipMemcpyDeviceToHost));
    if (val_host)
        NW_HIP(hipMemcpy(val_host, p->vals + z0, off[n] * sizeof(float), hipMemcpyDeviceToHost));
    return NW_OK;
}

int nw_electric_half_step(const nw_propagator* p, float half_dt_ms, float* v, const float* sum_g,
                          const float* drive_v, const float* sum_i, const float* inv_c,
                          float* integral_v, void* stream) {
    NW_REQUIRE(p && v && sum_g && drive_v && sum_i && inv_c, "nw_electric_half_step: null argument");
    HalfArgs a{p->cell_row0, p->cell_nz0, p->row_off, p->vals, p->cols, p->rowsum, p->order,
               p->max_cell_rows, half_dt_ms, v, sum_g, drive_v, sum_i, inv_c, integral_v, p->wgt};
    if (p->chebyshev) {
        NW_REQUIRE(p->max_cell_rows <= kChebBuf, "cell of %d segments exceeds the Chebyshev store",
                   p->max_cell_rows);
        // built for the store's product, the workgroup size and the largest cell
        auto launch = [&](auto kind) {
            constexpr int KD = decltype(kind)::value;
            const unsigned grid = (unsigned)p->num_cells;
            hipStream_t s = (hipStream_t)stream;
            if (p->cheb_small)
                k_half_step_cheb<1, kChebSmallBlock, KD><<<grid, kChebSmallBlock, 0, s>>>(
                    a, cheb_store(p));
            else if (p->cheb_maxj == 1)
                k_half_step_cheb<1, kChebBlock, KD><<<grid, kChebBlock, 0, s>>>(a, cheb_store(p));
            else
                k_half_step_cheb<kChebRows, kChebBlock, KD><<<grid, kChebBlock, 0, s>>>(
                    a, cheb_store(p));
        };
        if (p->rational) launch(std::integral_constant<int, kProdRational>{});
        else launch(std::integral_constant<int, kProdChebyshev>{});
        NW_HIP(hipGetLastError());
        return NW_OK;
    }
    const bool sym = p->symmetric != 0;
    // general: x and the product as FP32; symmetric: y as FP32 and z as FP64
    size_t lds = (sym ? sizeof(double) + sizeof(float) : 2 * sizeof(float)) *
                 (size_t)p->max_cell_rows;
    NW_REQUIRE(lds <= 160 * 1024 - 64, "cell of %d segments exceeds LDS", p->max_cell_rows);
    if (sym)
        k_half_step<true><<<(unsigned)p->num_cells, 256, lds, (hipStream_t)stream>>>(a);
    else
        k_half_step<false><<<(unsigned)p->num_cells, 256, lds, (hipStream_t)stream>>>(a);
    NW_HIP(hipGetLastError());
    return NW_OK;
}

}  // extern "C"
