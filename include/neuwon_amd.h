/*
 * neuwon_amd -- C ABI of the MI355X-native per-timestep advance of NEUWON.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative code on failure; the message of the
 *     last failure on the calling thread is returned by nw_last_error().  Nothing aborts.
 *   - Device pointers are plain addresses of device memory owned by the caller (the Python host
 *     allocates them as torch tensors); host pointers are named *_host.
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *   - Per-cell CSR propagators (nw_propagator) are owned by the library.
 *   - Calls that share device buffers on one stream are ordered by that stream; a handle is not
 *     thread-safe (the Python host serialises calls with a lock).
 *
 * Units follow the reference: mV, ms, Siemens, Farads, Amperes.
 */
#ifndef NEUWON_AMD_H
#define NEUWON_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define NW_OK 0
#define NW_ERR_ARG (-1)
#define NW_ERR_HIP (-2)
#define NW_ERR_BLAS (-3)
#define NW_ERR_ALLOC (-4)

/* Message of the last failure on this thread ("" if none). */
const char* nw_last_error(void);

/* Library version string, device name / CU count / LDS bytes per workgroup of `device`. */
const char* nw_version(void);
int nw_device_info(int device, char* name_out, int name_len, int* num_cus, int* lds_bytes);

/* ------------------------------------------------------------------------------------------
 * Propagator (replaces Electric._compute_propagator_matrix, neuwon/rxd/neuron/electric.py:79-117)
 *
 * Segments are grouped by cell (the reference sorts Segment by 'neuron', neuron.py:12, which makes
 * the global matrix block diagonal).  cell_ptr_host[c]..cell_ptr_host[c+1] are the segments of cell
 * c; parent_local_host holds each segment's parent as an index local to its cell, -1 for the root.
 * R_host (Ohm) and C_host (F) are the per-segment axial resistance and capacitance.
 * For every cell the library builds A*dt exactly as electric.py:92-108, takes expm in FP64 on the
 * device, drops entries below `prune` (electric.py:114-115), and keeps per-cell CSR rows with FP32
 * values and 16-bit local column indices (rows padded with zeros to a multiple of 4 entries).  The
 * diagonal is stored so that every FP32 row sums, in FP64, to the FP64 row sum of the pruned matrix.
 * ------------------------------------------------------------------------------------------ */
typedef struct nw_propagator nw_propagator;

#define NW_PROPAGATOR_SYMMETRIC 1
#define NW_PROPAGATOR_CHEBYSHEV 2
#define NW_PROPAGATOR_RATIONAL 4

typedef struct nw_propagator_stats {
    int64_t num_cells;
    int64_t num_rows;
    int64_t nnz;             /* entries kept by the pruning                                       */
    int64_t stored;          /* entries stored (FP32 value + u16 column), rows padded to 4        */
    int32_t max_cell_rows;
    int32_t max_row_nnz;
    double  build_seconds;
    int64_t device_bytes;    /* bytes held on the device by the handle                            */
    int32_t symmetric;       /* 1: symmetric format (see nw_propagator_build_ex)                  */
    int32_t chebyshev;       /* 1: no stored matrix, a Chebyshev series of the tree operator      */
    double  mean_terms;      /* Chebyshev store: series terms per row and product (row-weighted)  */
    int32_t workgroup_threads; /* Chebyshev store: threads per workgroup of its launches: 512 when
                                  every cell fits one row group of 512 slots, else 1024; 0 for the
                                  stored formats                                                  */
    int32_t row_groups;      /* Chebyshev store: row groups per thread its launches are built for */
    int32_t rational;        /* 1: the rational store (NW_PROPAGATOR_RATIONAL)                    */
    int32_t max_levels;      /* rational store: levels of the deepest cell's tree                 */
    double  mean_levels;     /* rational store: levels per cell, weighted by segments             */
} nw_propagator_stats;

int nw_propagator_build(int device, int64_t num_cells, const int64_t* cell_ptr_host,
                        const int32_t* parent_local_host, const double* R_host, const double* C_host,
                        double time_step_ms, double prune, nw_propagator** out);
/* flags = NW_PROPAGATOR_SYMMETRIC stores S = W P W^-1 (w = sqrt(C), symmetric because
 * A = C^-1 L with L symmetric, electric.py:92-108) on and above the diagonal only, halving
 * the stored entries; a pair (i, j) is kept when max(P_ij, P_ji) >= prune.
 * P x = W^-1 S W x.  Read by nw_hh_step, nw_hh_adjoint and nw_electric_half_step; nw_ca_step
 * refuses it. */
int nw_propagator_build_ex(int device, int64_t num_cells, const int64_t* cell_ptr_host,
                           const int32_t* parent_local_host, const double* R_host,
                           const double* C_host, double time_step_ms, double prune, int32_t flags,
                           nw_propagator** out);
/* flags = NW_PROPAGATOR_CHEBYSHEV stores no matrix: P x = exp(B) x with B = A dt (the tree
 * operator, 2(n-1) edge coefficients per cell) is evaluated as the Chebyshev series
 * sum_k c_k T_k(2 B / rho + I) x on rho = the spectral radius of B (B is similar to a symmetric
 * matrix; its largest eigenvalue by bisection with inertia counts, plus a 1e-8 margin), with
 * c_0 = e^(-rho/2) I_0(rho/2), c_k = 2 e^(-rho/2) I_k(rho/2), truncated where the remaining
 * coefficients fall below 1e-10 (nw_chebyshev_exp_coefficients).  Nothing is pruned: the
 * product is the reference's exp(A dt) x at its default threshold (electric.py:114-115, machine
 * epsilon) to about 1e-10 of the input's magnitude.  Cells whose rows have at most 3
 * neighbours run in depth-first preorder slots (the rows next to a row's slot from registers or
 * the neighbouring lane, two far neighbours from LDS); other cells keep up to 3 neighbours per
 * row in natural order (more in an overflow list); cells up to 5120 segments.  Read by
 * nw_hh_step (which also writes the adjoint's record), nw_electric_half_step and nw_hh_adjoint
 * (P^T y = C P (C^-1 y), the same series); nw_ca_step refuses it.
 * flags = NW_PROPAGATOR_RATIONAL stores no matrix either: exp(B) x is evaluated as
 * a0 x + 2 Re sum_j alpha_j (B - theta_j I)^-1 x, a type (10, 10) rational approximation of exp
 * on (-inf, 0] with five poles in the upper half plane (maximum error 3e-10; r(0) = 1 exactly, so
 * constant vectors are kept; nw_rational_coefficients).  Each shifted system is solved exactly
 * by elimination over the cell's tree: rows in level order (the root, its children, their
 * children, ...), a parent's children adjacent; leaves to root then root to leaves, one
 * workgroup barrier per level; the pivots 1 / d_i of every pole are built once in FP64 from the
 * FP32 edge coefficients (their row sums are zero) and read from memory.  Nothing is pruned (the
 * reference's default threshold, electric.py:114-115).  Cells up to 5120 segments, at most 63
 * children per segment, at most 4095 levels.  Read by nw_hh_step, nw_electric_half_step and
 * nw_hh_adjoint; nw_ca_step refuses it. */
int nw_propagator_stats_get(const nw_propagator* p, nw_propagator_stats* out);
/* Chebyshev store: the spectrum bound rho and series coefficients of one cell
 * (coef_host may be NULL: only *K_out, the index of the last kept coefficient, and *rho_out). */
int nw_propagator_cell_chebyshev(const nw_propagator* p, int64_t cell, int32_t* K_out,
                                 double* rho_out, double* coef_host);
/* Host routine: coefficients of exp((x/2)(t - 1)) = sum_k c_k T_k(t) for x = rho >= 0, i.e.
 * c_0 = e^(-x/2) I_0(x/2), c_k = 2 e^(-x/2) I_k(x/2) (modified Bessel functions by Miller's
 * backward recurrence, normalised by e^y = I_0(y) + 2 sum_k I_k(y)), up to the last k with
 * c_k > tol; the sum of the dropped ones is added to c_0, so sum_k c_k = 1 (a constant vector is
 * kept exactly).  Writes at most `capacity` values; returns the number of coefficients (K + 1), or
 * a negative error code. */
int nw_chebyshev_exp_coefficients(double rho, double tol, double* coef_host, int32_t capacity);
/* The rational store's approximation: out[0] = a0, then per pole theta_re, theta_im, alpha_re,
 * alpha_im (1 + 4 * 5 doubles).  No counterpart in the reference (it forms expm(A dt) with scipy,
 * electric.py:111-112). */
int nw_rational_coefficients(double* out, int32_t capacity);
/* Symmetric format only: the per-segment w of one cell. */
int nw_propagator_cell_weights(const nw_propagator* p, int64_t cell, double* w_host);
int nw_propagator_destroy(nw_propagator* p);

/* One cell's pruned FP64 propagator, built by the same device path, copied back as CSR.
 * Two calls: with val_host == NULL it only returns the entry count in *nnz_out. */
int nw_propagator_cell_f64(int device, int32_t n, const int32_t* parent_local_host,
                           const double* R_host, const double* C_host, double time_step_ms,
                           double prune, int64_t* nnz_out, int64_t* row_ptr_host,
                           int32_t* col_host, double* val_host);

/* Copy one cell's stored (FP32 value, u16 column) CSR back to the host (for checks).  Rows are
 * padded to a multiple of 4 entries (value 0, column = the row); within each 256-entry chunk of
 * m entries the k-th entry in column order sits at slot 4 (k mod m/4) + k / (m/4). */
int nw_propagator_cell_export(const nw_propagator* p, int64_t cell, uint32_t* row_off_host,
                              uint16_t* col_host, float* val_host);

/* ------------------------------------------------------------------------------------------
 * Electric half step (replaces Electric._advance_electric, electric.py:57-76)
 *   dv = dt*I/C ; alpha = exp(-dt*1e-3*g/C) ; x = d - (d - v)*alpha + dv ; v <- P x
 * All arrays are FP32 per segment, in the cell-grouped order used to build `p`.
 * integral_v may be NULL (the reference writes it, electric.py:76, but nothing reads it).
 * ------------------------------------------------------------------------------------------ */
int nw_electric_half_step(const nw_propagator* p, float half_dt_ms, float* v, const float* sum_g,
                          const float* drive_v, const float* sum_i, const float* inv_c,
                          float* integral_v, void* stream);

/* ------------------------------------------------------------------------------------------
 * Accumulation (replaces _SpeciesType._apply_accumulators / _NonspecificConductance, species.py:
 * 147-160, 252-260, and rxd_model.py:93-109): segmented reduction over instances sorted by
 * segment.  seg_inst_ptr[s]..seg_inst_ptr[s+1] are the instances on segment s (CSR offsets, so no
 * atomics and a deterministic order).  Every instance carries a conductance and a reversal
 * potential; instance k's conductance is g_inst[g_index[k]] (g_index may be NULL: g_inst[k]), so
 * sources written by different mechanisms into one buffer are read in segment order.
 * nonspec_i (may be NULL) is the per-segment nonspecific current.  Sums are formed in FP64.
 *   sum_g[s] = sum g ; drive_v[s] = sum(g E) / sum_g (0 where sum_g == 0) ; sum_i[s] = nonspec_i
 * ------------------------------------------------------------------------------------------ */
int nw_accumulate(int64_t num_segments, const int64_t* seg_inst_ptr, const float* g_inst,
                  const int64_t* g_index, const float* e_inst, const float* nonspec_i,
                  float* sum_g, float* drive_v, float* sum_i, void* stream);

/* AP detector of every neuron (replaces Neuron._advance_AP_detector, neuron.py:72-76):
 *   over = v[seg[i]] >= threshold; detected[i] = over && !state[i]; state[i] = over. */
int nw_ap_detect(int64_t num, const int32_t* seg, const float* v, float threshold,
                 uint8_t* state, uint8_t* detected, void* stream);

/* ------------------------------------------------------------------------------------------
 * LTI kinetic advance (replaces the kernel generated by lti_sim/codegen.py:39-106)
 *   node_table: [num_nodes][S][S] FP32 propagators (row-major: entry [node][row][col]) on a linear
 *   voltage grid vmin..vmax with num_nodes-1 buckets; linear interpolation between nodes.
 *   states: S arrays of M FP32 (SoA, array s at states + s*state_stride).
 *   voltage is gathered through seg_index (may be NULL: instance i reads voltage[i]).
 *   conserve_sum <= 0 disables the CONSERVE renormalisation.
 *   n_steps > 1 repeats the advance with the inputs held constant (every step identical to one
 *   call with n_steps = 1).
 * If cond_out != NULL, cond_out[i] = gmax[i] * sum_s weights[s] * state_s  (the breakpoint of a
 * kinetic mechanism whose conductance is a weighted sum of states, e.g. ampa13.mod BREAKPOINT).
 * ------------------------------------------------------------------------------------------ */
int nw_lti_advance(int32_t S, int32_t num_nodes, float vmin, float vmax, const float* node_table,
                   int64_t M, const float* voltage, const int32_t* seg_index, float* states,
                   int64_t state_stride, float conserve_sum, int32_t n_steps, const float* weights,
                   const float* gmax, float* cond_out, void* stream);

/* ------------------------------------------------------------------------------------------
 * LTI kinetic tables on the device and the advance grouped by voltage bucket (the table upload
 * and kernel of lti_sim/codegen.py:29-106 for schemes whose one input is the voltage; table
 * layout and lookup as nw_lti_advance).
 * nw_lti_table_create keeps the caller's FP32 node_table (device, [num_nodes][S][S], it must
 * outlive the handle) and builds two device copies of it: the values widened to FP64 and an FP32
 * copy with every node padded to a multiple of 4 values.
 * nw_lti_table_advance advances M instances one step (n_steps = 1): per instance
 *   y0 = T_b x, y1 = T_(b+1) x, x' = y0 + f (y1 - y0)   (FP64, each row summed over the columns in
 *   ascending order), then CONSERVE and the conductance output as nw_lti_advance.
 * A wavefront whose instances fall into at most group_max distinct buckets (-1: the library's
 * default, 2) reads the two node matrices of each bucket once, by scalar loads; other
 * wavefronts read every instance's nodes from the padded copy.  Both forms give the same result
 * bit for bit.  n_steps > 1 holds the inputs over the launch and runs nw_lti_advance.
 * ------------------------------------------------------------------------------------------ */
typedef struct nw_lti_table nw_lti_table;

/* Table build of lti_sim/lti_model.py:51-65 on the device: P[m] = expm(A[m] * time_step) for
 * `count` S x S FP64 rate matrices (device, row-major, S <= 16), every column of P[m] divided by
 * its sum when normalise_columns (the reference normalises, lti_model.py:63-64).  One wavefront
 * per matrix on the FP64 matrix cores: X = A dt + lambda I (lambda = max |A_ii dt|, non-negative
 * for a rate matrix), scaled by 2^-s to a 1-norm <= 1/2, an 18-term Taylor series, s squarings of
 * e^(-lambda 2^-s) expm(X 2^-s); with normalise_columns and columns of A that sum to zero (to
 * their rounding), every squaring's columns are rescaled to sum one (the propagators of a
 * conserving scheme are column-stochastic at every stage). */
int nw_lti_expm_batched(int64_t count, int32_t S, const double* A, double time_step,
                        int32_t normalise_columns, double* P, void* stream);

int nw_lti_table_create(int device, int32_t S, int32_t num_nodes, float vmin, float vmax,
                        const float* node_table, nw_lti_table** out);
int nw_lti_table_destroy(nw_lti_table* t);
int nw_lti_table_advance(const nw_lti_table* t, int64_t M, const float* voltage,
                         const int32_t* seg_index, float* states, int64_t state_stride,
                         float conserve_sum, int32_t n_steps, const float* weights,
                         const float* gmax, float* cond_out, int32_t group_max, void* stream);

/* ------------------------------------------------------------------------------------------
 * LTI kinetic advance with one or two inputs of either mapping (replaces the kernel generated by
 * lti_sim/codegen.py:39-106 for schemes whose inputs are the voltage and/or a concentration, e.g.
 * AMPA.mod's glutamate, NMDA.mod's transmitter and voltage; inputs mapped as
 * lti_sim/inputs.py:72-109).  Input k of instance i is x = values[index ? index[i] : i] and
 * maps to a fractional node position u:
 *   NW_INPUT_LINEAR:  u = (x - minimum) * (num_nodes - 1) / (maximum - minimum)
 *   NW_INPUT_LOG2:    u = (log2(x + scale) - log2(minimum + scale)) * (num_nodes - 1)
 *                         / (log2(maximum + scale) - log2(minimum + scale))
 * node b = trunc(u) clamped to [0, num_nodes - 2] (fraction 1 above, 0 below).  node_table holds
 * [node_0][node_1][S][S] FP32 propagators (row-major entries, input 0 slowest); one input
 * interpolates linearly, two bilinearly.  States, CONSERVE and the conductance output as in
 * nw_lti_advance.  The inputs are read once per launch (held over n_steps).
 * ------------------------------------------------------------------------------------------ */
#define NW_INPUT_LINEAR 0
#define NW_INPUT_LOG2 1

typedef struct nw_lti_input {
    const float* values;    /* device */
    const int32_t* index;   /* device, may be NULL */
    int32_t mapping;        /* NW_INPUT_* */
    int32_t num_nodes;      /* >= 2 */
    double minimum, maximum;
    double scale;           /* NW_INPUT_LOG2 */
} nw_lti_input;

int nw_lti_advance_nd(int32_t S, int32_t num_inputs, const nw_lti_input* inputs_host,
                      const float* node_table, int64_t M, float* states, int64_t state_stride,
                      float conserve_sum, int32_t n_steps, const float* weights, const float* gmax,
                      float* cond_out, void* stream);

/* ------------------------------------------------------------------------------------------
 * HH mechanism instances (replaces the advance generated for hh.mod, nmodl/__init__.py:408-432
 * with the LTI or exact gate update, and its breakpoint, hh.mod:41-46): M instances, instance i on
 * segment seg_index[i] (NULL: segment i).  gate_mode 0: two-state LTI tables (layout as nw_hh_step);
 * 1: exact update with rate factor q10.  Per instance gbar*magnitude arrays g_na, g_k, g_l (S).
 * g_out[3*i + {0,1,2}] receives the na, k and l conductances for nw_accumulate.
 * ------------------------------------------------------------------------------------------ */
int nw_hh_mechanism(int64_t M, int32_t gate_mode, float dt_ms, float q10, const float* gate_table,
                    int32_t num_nodes, float vmin, float vmax, const float* voltage,
                    const int32_t* seg_index, float* m, float* h, float* n, const float* g_na,
                    const float* g_k, const float* g_l, float* g_out, void* stream);

/* ------------------------------------------------------------------------------------------
 * Fused staggered step for cells carrying leak + HH Na (m^3 h) + HH K (n^4), every gate a two
 * state kinetic scheme advanced by its LTI table (replaces RxD_Model.advance, rxd_model.py:63-123,
 * for this mechanism set).  One workgroup advances one cell through n_steps steps:
 *   half step (dt/2) -> mechanisms (dt, voltage at t+dt/2) -> half step (dt/2) -> AP detector.
 * gate_table: 3 gates (m, h, n) x num_nodes x {a, b}: open' = a(v) + b(v)*open, linear
 *   interpolation between nodes of the vmin..vmax grid (a = P_OC, b = P_OO - P_OC of the 2x2
 *   column-normalised expm, lti_model.py:51-65).
 * Per segment FP32 arrays: v, m, h, n, g_na, g_k, g_l (Siemens = gbar*area), inv_c (1/F).
 * Stimulus: per cell one segment (stim_local) receives stim_amp_a * k amperes on input tick t, where
 *   k = 1 for a constant clamp (stim_mode 1) or the number of Poisson pulse onsets within the last
 *   pulse_ticks ticks (stim_mode 2; onset on tick t iff hash(seed, cell, t) < poisson_threshold).
 * gate_mode 1 instead advances each gate with the exact solution of hh.mod's DERIVATIVE block
 *   (x' = inf + (x - inf) exp(-(alpha+beta) dt), nmodl/solver.py:48-55; rates hh.mod:62-94).
 * conductances_valid = 0 reproduces the reference's first step, whose first half step sees zero
 *   accumulated conductance.
 * AP detector (neuron.py:72-76) on segment ap_local of each cell: ap_state (u8 per cell) holds
 *   the over-threshold flag, ap_count (u32 per cell) counts detections.
 * trace (may be NULL): trace[cell * trace_stride + step] = v of segment ap_local after the step.
 * record (may be NULL): see nw_hh_adjoint; (5 n_steps + 4) arrays of num_rows floats.
 * ------------------------------------------------------------------------------------------ */
typedef struct nw_hh_params {
    float   dt_ms;
    float   vmin, vmax;
    int32_t num_nodes;
    float   e_na, e_k, e_l;
    int32_t stim_mode;        /* 0 none, 1 constant, 2 Poisson pulses                              */
    float   stim_amp_a;
    uint32_t stim_seed;
    uint32_t poisson_threshold;
    int32_t pulse_ticks;
    float   ap_threshold;
    int32_t conductances_valid;
    int64_t step0;            /* global step index of the first step (input ticks continue)      */
    int32_t block_threads;    /* 128/256/512/1024; 0 = 512 (symmetric format) or 256 (general)    */
    int32_t gate_mode;        /* 0: two-state LTI tables; 1: exact exponential update of hh.mod    */
    float   q10;              /* gate_mode 1: rate factor 3^((celsius-6.3)/10), hh.mod:67         */
    int32_t launch_mode;      /* NW_LAUNCH_* (matrix-free stores; the stored formats ignore it)    */
    int32_t cheb_launch;      /* NW_CHEB_LAUNCH_* (the Chebyshev store; the others ignore it)      */
    int32_t reserved;         /* 0                                                                 */
} nw_hh_params;

/* cheb_launch of nw_hh_step on the Chebyshev store: how the steps map to kernel launches (every
 * form does the same arithmetic in the same order, so the results are bitwise the same):
 *   NW_CHEB_LAUNCH_DEFAULT    two steps per launch when the store is built for five row groups
 *                             per thread, one step per launch when built for one;
 *   NW_CHEB_LAUNCH_HALF       one launch per half step;
 *   NW_CHEB_LAUNCH_STEP       one launch per step (a cell's workgroup runs both half steps);
 *   NW_CHEB_LAUNCH_TWO_STEPS  one launch per two steps (an odd n_steps' last step alone). */
#define NW_CHEB_LAUNCH_DEFAULT 0
#define NW_CHEB_LAUNCH_HALF 1
#define NW_CHEB_LAUNCH_STEP 2
#define NW_CHEB_LAUNCH_TWO_STEPS 3

/* launch_mode of nw_hh_step on the matrix-free (Chebyshev, rational) stores, which run their
 * launches in the form cheb_launch gives (the Chebyshev store) or one per half step (the rational
 * store):
 *   NW_LAUNCH_DIRECT         the launches are issued one by one on `stream`;
 *   NW_LAUNCH_GRAPH          they are replayed as one HIP graph launch on `stream`: a graph is
 *                            captured (on a stream of the handle's own) the first time an argument
 *                            set is seen -- every argument but prm->step0, which the kernels read
 *                            from device memory written just before each replay -- and kept with
 *                            the handle (the last 4 argument sets), so repeated calls replay it;
 *                            on a stream that is itself being captured the launches are direct;
 *   NW_LAUNCH_GRAPH_PREPARE  the graph of this argument set is captured and kept, nothing runs
 *                            (e.g. before a timed region).
 * A graph holds the array addresses it was captured with: calls with other arrays capture again. */
#define NW_LAUNCH_DIRECT 0
#define NW_LAUNCH_GRAPH 1
#define NW_LAUNCH_GRAPH_PREPARE 2

int nw_hh_step(const nw_propagator* p, const nw_hh_params* prm, int32_t n_steps,
               const float* gate_table, float* v, float* m, float* h, float* n, const float* g_na,
               const float* g_k, const float* g_l, const float* inv_c, const int32_t* stim_local,
               const int32_t* ap_local, uint8_t* ap_state, uint32_t* ap_count, float* trace,
               int64_t trace_stride, float* record, void* stream);

/* ------------------------------------------------------------------------------------------
 * Reverse pass of the fused step (no counterpart in the reference, which has no differentiation:
 * it is the adjoint of nw_hh_step with table gates and a constant stimulus).  `record` is the
 * record nw_hh_step wrote for the same n_steps (num_rows-long arrays v, m, h, n at the start of
 * every step and v after its first half step, 5 per step, then v, m, h, n after the last step),
 * with any store; `p` may be any store too (the Chebyshev store applies P^T y = C P (C^-1 y) as
 * the same series, one launch per half step).
 * Loss: sum over the steps of (v[soma_local[cell]] after the step - target[cell*stride + step])^2.
 * lam_v/lam_m/lam_h/lam_n (FP32, per segment) carry the adjoint of the state from the following
 * launch (zeros before the last one); grad_na/k/l (FP64, per segment) and grad_theta (FP64,
 * [cell][6]: d/dt_alpha, d/dt_beta for m, h, n) and loss (FP64 per cell) are accumulated.
 * table6: [3 gates][num_nodes][6] = a, b, da/dt_alpha, db/dt_alpha, da/dt_beta, db/dt_beta.
 * ------------------------------------------------------------------------------------------ */
int nw_hh_adjoint(const nw_propagator* p, const nw_hh_params* prm, int32_t n_steps,
                  const float* record, const float* table6, const float* g_na, const float* g_k,
                  const float* g_l, const float* inv_c, const int32_t* soma_local,
                  const float* target, int64_t target_stride, float* lam_v, float* lam_m,
                  float* lam_h, float* lam_n, double* grad_na, double* grad_k, double* grad_l,
                  double* grad_theta, double* loss, void* stream);

/* ------------------------------------------------------------------------------------------
 * Intracellular Ca2+ with a first-order buffer and diffusion (replaces _SpeciesInstance._advance,
 * species.py:54-70, whose diffusion matrix is a stub at species.py:74-75, together with a buffer
 * mechanism writing the Ca derivative accumulator, nmodl/__init__.py:204-207).  `p` is the
 * diffusion propagator of the half step (built by nw_propagator_build with R = 1e-3/k,
 * k = D*xarea/L per edge, and C = volume).  Per step and cell:
 *   c = P max(0, c + dcdt dt/2);  CaB' = inf + (CaB - inf) exp(-(kon c + koff) dt),
 *   inf = btot kon c/(kon c + koff);  dcdt = -(CaB' - CaB)/dt;  c = P max(0, c + dcdt dt/2).
 * Arrays FP32 per compartment (any concentration unit consistent with kon); `cab` holds the bound
 * buffer minus cab_ref (so its small per-step changes keep FP32 precision next to free Ca).
 * ------------------------------------------------------------------------------------------ */
typedef struct nw_ca_params {
    float   dt_ms;
    float   kon;      /* per concentration unit per ms */
    float   koff;     /* per ms */
    float   btot;     /* total buffer, concentration unit */
    double  cab_ref;  /* reference level of the bound buffer */
    int32_t block_threads;
    int32_t pad_;
} nw_ca_params;

int nw_ca_step(const nw_propagator* p, const nw_ca_params* prm, int32_t n_steps, float* ca,
               float* cab, float* dcdt, void* stream);

/* partials[b] = FP64 sum over block b's share of w[i] * (x[i] + y[i] + z_scale * z[i]) (y, z may
 * be NULL); the caller sums the num_partials values (mass conservation checks). */
int nw_weighted_sum(int64_t n, const double* w, const float* x, const float* y, const float* z,
                    double z_scale, double* partials, int32_t num_partials, void* stream);

/* nw_weighted_sum without a workspace argument.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The num_partials partials (at most NW_SUM_SCRATCH_PARTIALS) live in a device
 * array of the library's own, one per device, held until the process ends; a second kernel sums
 * them in order into total[0] (FP64, device), the value a host loop over nw_weighted_sum's
 * partials gives.  The call allocates nothing.  Calls that overlap on one device (two streams,
 * two threads) share that array and race. */
#define NW_SUM_SCRATCH_PARTIALS 4096
int nw_weighted_sum_total(int64_t n, const double* w, const float* x, const float* y,
                          const float* z, double z_scale, int32_t num_partials, double* total,
                          void* stream);

/* Stimulus generator, the same hash the step kernel evaluates (exposed for checks). */
uint32_t nw_stim_hash(uint32_t seed, uint32_t cell, uint32_t tick);

/* ------------------------------------------------------------------------------------------
 * Clock events on the device (replaces the per-tick callbacks TimeSeries._play_implementation,
 * TimeSeries._record_implementation and Trace._callback, neuwon/database/time.py:224-232,
 * 287-302, 509-515, that Clock.tick runs, time.py:114-117).  One launch per tick of one clock:
 * t = *tick + 1 is the clock's count after the tick, and *tick holds t when the launch ends, so a
 * captured graph replays successive ticks.  *num_events is read from device memory as well.
 * Events act on ticks first_tick <= t < end_tick, in table order (the order of registration):
 *   NW_EVENT_SET / NW_EVENT_ADD: k = first_index + (t - first_tick), modulo length when loop != 0;
 *     column[row] = wave[k] (SET) or (float)(wave[k] + (double)column[row]) (ADD), FP32 columns.
 *   NW_EVENT_RECORD: ring[slot * ring_capacity + t % ring_capacity] = column[row] (as FP32).
 *   NW_EVENT_TRACE: state[0..2] = mean, var, start of the exponentially weighted statistics of
 *     column[row] (time.py:509-515): FP64 with every product and sum rounded separately, as the
 *     reference's Python float arithmetic rounds them.
 * A run of consecutive RECORD/TRACE events (indices i .. run_end-1) is sampled in parallel; SET and
 * ADD events are applied one at a time, so several players on one element compose in order.
 * ------------------------------------------------------------------------------------------ */
#define NW_EVENT_SET 0
#define NW_EVENT_ADD 1
#define NW_EVENT_RECORD 2
#define NW_EVENT_TRACE 3
#define NW_DTYPE_F32 0
#define NW_DTYPE_U8 1
#define NW_DTYPE_I32 2

typedef struct nw_event {
    int32_t kind;          /* NW_EVENT_* */
    int32_t dtype;         /* NW_DTYPE_* of the column (SET/ADD: FP32 only) */
    void* column;          /* device column */
    int64_t row;           /* element of the column */
    int64_t first_tick;
    int64_t end_tick;
    const double* wave;    /* SET/ADD: samples (device, FP64) */
    int64_t length;        /* SET/ADD: number of samples */
    int64_t first_index;   /* SET/ADD: sample written on first_tick */
    int64_t slot;          /* RECORD: row of the ring */
    double* state;         /* TRACE: mean, var, start (device, FP64) */
    double alpha, beta;    /* TRACE: exp(-dt / period), 1 - alpha */
    int32_t loop;          /* SET/ADD: wrap around */
    int32_t run_end;       /* RECORD/TRACE: end of this run of samplers */
} nw_event;

int nw_clock_tick(int64_t* tick, const int32_t* num_events, const nw_event* events, float* ring,
                  int64_t ring_capacity, void* stream);

/* TraceAll._callback (time.py:627-637) for every instance of an FP32 column, FP64 arithmetic:
 *   diff = x - mean; incr = beta diff; mean += incr; var = alpha (var + diff incr).
 * var may be NULL (no variance). */
int nw_trace_all(int64_t n, const float* x, float* mean, float* var, double alpha, double beta,
                 void* stream);

#ifdef __cplusplus
}
#endif
#endif
