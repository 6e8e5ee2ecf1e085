/*
 * born.h -- C ABI of the MI355X Born scattering operator pair.
 *
 * Replaces the CUDA MEX gateway `combined_quick_gpu`
 * (+scattering/combined_quick_gpu.cu:66-1094, documented in
 * +scattering/combined_quick_gpu.m:1-41), which MATLAB calls as
 *
 *     output = scattering.combined_quick_gpu( operator_born, mode, input )
 *
 * from operator_born.forward_scalar (operator_born.m:217) and
 * operator_born.adjoint_scalar (operator_born.m:286).  The MEX pulls its
 * tables out of the operator object on every call and frees everything at
 * the end (cu:201-391, :817-823, :1011-1088).  This library splits that into
 * a persistent plan that owns device-resident tables, and apply calls.
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) pairs, as in MATLAB's
 *     interleaved complex API (R2018a, combined_quick_gpu.m:31-32);
 *   - matrices are column-major, as MATLAB stores them;
 *   - index arrays are 1-based doubles exactly as MATLAB holds them;
 *   - observations are ordered measurement-major, then mix, then frequency
 *     (cu:773, :978);
 *   - a plan is not re-entrant: calls on one plan must not overlap.  Use one
 *     plan per stream.
 */
#ifndef BORN_H
#define BORN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; born_status_id() maps each to the reference's error
 * identifier string (cu:201, :228, :372, :385, :397, :414, :421, :428;
 * h:31-44 for the runtime error). */
enum born_status {
    BORN_OK = 0,
    BORN_ERR_NUMBER_ARGUMENTS = 1,       /* combined_quick_gpu:NumberArguments */
    BORN_ERR_NO_SYMMETRIC_GRIDS_SETUP = 2,/* combined_quick_gpu:NoSymmetricGridsSetup */
    BORN_ERR_NO_GPU_ACTIVE = 3,          /* combined_quick_gpu:NoScatteringOptionsGPUActive */
    BORN_ERR_NO_COMPLEX_DOUBLES = 4,     /* combined_quick_gpu:NoComplexDoubles */
    BORN_ERR_NO_OPERATOR_BORN = 5,       /* combined_quick_gpu:NoOperatorBorn */
    BORN_ERR_NO_NUMERIC_SCALAR_MODE = 6, /* combined_quick_gpu:NoNumericScalarMode */
    BORN_ERR_NO_VALID_DOUBLE_MATRIX = 7, /* combined_quick_gpu:NoValidDoubleMatrix */
    BORN_ERR_CONVERSION_FAILED = 8,      /* combined_quick_gpu:ConversionFailed */
    BORN_ERR_DEVICE = 9,                 /* adjoint_quick_gpu:ErrorCUDA */
    BORN_ERR_INVALID_ARGUMENT = 10,      /* born:InvalidArgument (bad sizes/indices) */
    BORN_ERR_UNSUPPORTED = 11,           /* born:Unsupported */
    BORN_ERR_OUT_OF_MEMORY = 12          /* born:OutOfMemory */
};

/* Numeric form of the grid contraction. */
enum born_precision {
    BORN_PRECISION_FP32 = 0,     /* plain FP32 operands, FP32 MFMA products      */
    BORN_PRECISION_SPLIT_BF16 = 1,/* hi+lo BF16 operands, FP32 accumulation       */
    BORN_PRECISION_SPLIT_FP16 = 2 /* hi+lo FP16 operands (scaled), FP32 accum.   */
};

typedef struct born_plan_s* born_plan_t;

/* One mixed voltage signal k of measurement m: settings(m).rx(k) and
 * prefactors{m}(k) (cu:298, :309-333, :352). */
typedef struct {
    int32_t N_f;                         /* frequencies of the mix (cu:298)                 */
    const double* indices_f_to_unique;   /* N_f, 1-based into the measurement's unique freqs
                                            (settings(m).indices_f_to_unique{k}, cu:317)    */
    int32_t N_active;                    /* active elements (cu:309-310)                    */
    const double* indices_active;        /* N_active, 1-based element indices               */
    const double* prefactors;            /* complex N_f x N_active, col-major (cu:352)      */
} born_mix_desc;

/* One selected sequential pulse-echo measurement m. */
typedef struct {
    int32_t N_f_unique;                  /* unique frequencies of the measurement           */
    const double* indices_f_to_unique;   /* N_f_unique, 1-based into the global unique freqs
                                            (sequence.indices_f_to_unique{m}, cu:278-281)   */
    const double* p_incident;            /* complex N_f_unique x N_points, col-major
                                            (incident_waves(m).p_incident, cu:817-823);
                                            NULL when the plan synthesizes it on the device */
    int32_t N_mix;
    const born_mix_desc* mixes;
} born_measurement_desc;

/* The tables the MEX reads from a scattering.operator_born object
 * (cu:201-391), plus the FOV grid shape (FOV.shape.grid.N_points_axis) and
 * the array shape, which the plan uses to verify the lateral structure of
 * the shift map. */
typedef struct {
    int32_t N_dimensions;                /* FOV grid dimensions: 2 or 3                     */
    int32_t N_points_axis[3];            /* (N_x, N_z, -) or (N_x, N_y, N_z)               */
    int32_t N_elements_axis[2];          /* (N_ex, 1) or (N_ex, N_ey)                       */
    int32_t N_points;                    /* sequence.size(2) (cu:222)                       */
    int32_t N_elements;                  /* xdc_array.N_elements (cu:234)                   */
    int32_t N_f_unique;                  /* rows of h_ref (cu:391)                          */
    const double* h_ref;                 /* complex N_f_unique x N_points (cu:382-391)      */
    const double* indices_grid_FOV_shift;/* N_points x N_elements, 1-based (cu:231)         */
    int32_t N_measurements;              /* numel(indices_measurement_sel) (cu:241)         */
    const born_measurement_desc* measurements;
    int32_t device;                      /* options.momentary.gpu.index (cu:378)            */
    int32_t precision;                   /* enum born_precision                             */
} born_tables_desc;

/* Geometry for the device-side table build: processing.greens_function
 * (greens_function.m:36-122) integrated over the reference element's face
 * (setup.m:1216-1281), the anti-aliasing taper (setup.m:916-1032), and the
 * incident field p_in = sum_e h_tx(shift_e) v_d(e) (sequence.m:768-861). */
typedef struct {
    int32_t N_dimensions;                /* 2 or 3                                          */
    int32_t N_points_axis[3];
    double delta_axis[3];                /* grid spacings                                   */
    double z0;                           /* first axial grid position                       */
    int32_t N_elements_axis[2];
    double pitch_axis[2];
    int32_t N_face_points_axis[2];       /* face integration points (grid_numbers)          */
    double face_delta_axis[2];           /* face grid spacings                              */
    const double* face_weights;          /* complex N_face x N_f_unique or NULL (= 1)
                                            (face.m:359-376)                                */
    int32_t N_f_unique;
    const double* frequencies;           /* Hz, global unique axis                          */
    const double* wavenumbers;           /* complex N_f_unique k-tilde (time_causal.m:99-155)*/
    int32_t filter_rx_kind, filter_tx_kind; /* 0 off, 1 raised cosine, 2 logistic, 3 boxcar */
    double filter_rx_param, filter_tx_param;
    int32_t N_measurements;
    const born_measurement_desc* measurements; /* p_incident must be NULL                    */
    const int32_t* N_tx_active;          /* per measurement                                 */
    const double* const* tx_indices_active; /* per measurement, 1-based                     */
    const double* const* v_d;            /* per measurement, complex N_f_unique(m) x N_tx_active
                                            (settings(m).v_d_unique, tx.m:264-286)          */
    int32_t device;
    int32_t precision;
} born_geometry_desc;

/* Plan creation from raw operator tables (the MEX path).  Converts the
 * FP64 tables to the device layout once (the reference re-converts and
 * re-uploads on every call, cu:548-593, :817-847). */
int born_plan_create_tables(const born_tables_desc* desc, born_plan_t* plan);

/* Plan creation from geometry: builds h_ref (FP64 evaluation, FP32 store)
 * and p_incident on the device. */
int born_plan_create_geometry(const born_geometry_desc* desc, born_plan_t* plan);

int born_plan_destroy(born_plan_t plan);

/* op_A( x, 0 ): size of the operator [N_observations, N_points]
 * (operator.m:1035-1046). */
int born_plan_size(born_plan_t plan, int64_t* N_observations, int64_t* N_points);

/* Select the numeric form used by later apply calls (enum born_precision). */
int born_plan_set_precision(born_plan_t plan, int32_t precision);

/* combined_quick_gpu( op, 1, gamma ): host FP64 in/out, MEX semantics --
 * forward occupancy rule (cu:466-488), zero output when nothing is occupied
 * (cu:496); with few occupied points the gather path of
 * born_plan_set_sparse_forward.  gamma: complex N_points x N_rhs; u:
 * complex N_obs x N_rhs. */
int born_forward(born_plan_t plan, const double* gamma, int32_t N_rhs, double* u);

/* combined_quick_gpu( op, 2, u ): host FP64 in/out; all N_points rows
 * (cu:460, :999-1003). */
int born_adjoint(born_plan_t plan, const double* u, int32_t N_rhs, double* gamma_hat);

/* Device variants for PyTorch callers: complex64 device pointers in the same
 * column-major layout, enqueued on `stream` (a hipStream_t; NULL = default
 * stream).  Asynchronous with respect to the host. */
int born_forward_dev(born_plan_t plan, const void* gamma_c64, int32_t N_rhs, void* u_c64, void* stream);
int born_adjoint_dev(born_plan_t plan, const void* u_c64, int32_t N_rhs, void* gamma_hat_c64, void* stream);

/* The operator pair applied without allocating: mode 1 is the forward of
 * born_forward_dev, any other value the adjoint of born_adjoint_dev (the
 * MEX's mode rule, cu:404-409), same buffers and layouts, computed by the
 * same kernels.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, serving
 * loops, harnesses that forbid it.  It calls no hipMalloc or hipFree:
 * instead of the plan's per-rhs workspace it uses a one-rhs layout built at
 * plan creation and a static device array of BORN_NOALLOC_SCRATCH_BYTES that
 * the library holds on every device it runs on until the process ends.  The
 * right-hand sides run one after another.  Every call on a device shares
 * that array, so calls that overlap on one device (other streams, other
 * threads, other plans) race and must be serialized by the caller.  A plan
 * off the structured path, or whose one-rhs layout needs more scratch than
 * the array holds, gets BORN_ERR_UNSUPPORTED with the sizes in
 * born_last_error().  Event timing (born_plan_set_timing) does not cover
 * these calls.  Results agree with born_forward_dev / born_adjoint_dev within
 * FP32 rounding (the forward's partial sums may be blocked differently). */
#define BORN_NOALLOC_SCRATCH_BYTES (64ull << 20)
int born_apply_dev_noalloc(born_plan_t plan, int32_t mode, const void* in_c64, int32_t N_rhs, void* out_c64,
                           void* stream);

/* Received energies of one pulse-echo measurement, canonical basis, no TGC:
 * energy_rx_canonical( op ) (operator_born.m:718-859) -- for every grid
 * point p, sum over the measurement's mixes and their frequencies of
 * |sum_a pref(f, a) h_ref(f, shift(p, e_a)) p_incident(f, p)|^2, i.e. the
 * squared norm of column p of the measurement's rows of A.  `measurement` is
 * the 0-based index among the plan's measurements; the reference requires a
 * single selected measurement (energy_rx_canonical:NoSingleMeasurementIndex).
 * Host variant: energy is N_points doubles.  Device variant: N_points FP64
 * on the device, enqueued on `stream`. */
int born_energy_rx(born_plan_t plan, int32_t measurement, double* energy);
int born_energy_rx_dev(born_plan_t plan, int32_t measurement, void* energy_f64, void* stream);

/* Time-domain mixed voltage signals: for every right-hand side and mix
 * signal (measurement-major, the order of the observations) the real signal
 * N_t * ifft(X, N_t, 'symmetric') of its zero-padded Fourier coefficients
 * X[q] = u(obs(mix, f)), q = f / delta_f -- format_voltages then
 * signal_matrix.signal (operator.m:900-1007; signal_matrix.m:293-397) with
 * time-axis lower bound 0.  `frequencies` are the plan's N_f_unique global
 * unique frequencies in Hz; delta_f <= 0 takes the smallest gap between
 * them; delta_t <= 0 takes 1 / (2 q_max delta_f) (signal_matrix.m:333), so
 * N_t = round(1 / (delta_f delta_t)).  u: complex64 device N_rhs x N_obs;
 * rf: float32 device N_rhs x N_mix_signals x N_t.  Synchronous on `stream`.
 * A mix must not repeat a frequency (its axis is increasing in the
 * reference), else BORN_ERR_UNSUPPORTED.  No counterpart in the MEX (a
 * MATLAB-side format). */
int born_time_signals(born_plan_t plan, const double* frequencies, double delta_f, double delta_t,
                      const void* u_c64, int32_t N_rhs, void* rf_f32, int64_t capacity, int32_t* N_t,
                      void* stream);

/* Forward with few occupied points (host entry point born_forward): the MEX
 * compacts the input to its occupied points (cu:454-499); with few of them
 * the plan computes u = sum_j A[:, p_j] gamma_j by gathers from the tables
 * (FP64 sums) instead of the full contraction.  max_points < 0: the plan
 * decides by estimated cost (default); 0: never; k > 0: whenever at most k
 * points are occupied.  Results agree with the contraction within its
 * tolerance.  No counterpart in the MEX. */
int born_plan_set_sparse_forward(born_plan_t plan, int32_t max_points);

/* Columns of A at the given grid points: the forward of unit vectors, which
 * the solvers use to extract columns (omp.m:177-184; cosamp.m:205-223
 * applies the operator to batches of up to 1024 unit vectors).  Gathered
 * from the tables directly, O(N_observations) per column, instead of one
 * full forward per unit vector.  points: n 1-based doubles.  Host variant:
 * out complex FP64 N_obs x n, column-major.  Device variant: complex64
 * n x N_obs on the device (the layout of born_forward_dev), complete on
 * return.  No counterpart in the MEX. */
int born_columns(born_plan_t plan, const double* points, int32_t n, double* out);
int born_columns_dev(born_plan_t plan, const double* points, int32_t n, void* out_c64, void* stream);

/* Time gain compensation (TGC), the linear transform forward_scalar applies
 * to u after the operator and adjoint_scalar applies before it
 * (operator_born.m:222-236, :263-277).  regularization.tgc.exponential
 * builds it as one discrete convolution per mix along the mix's Fourier
 * coefficients, cut off to the mix's length (tgc/exponential.m:118-213;
 * linear_transforms.convolutions.fft, fft.m:85-126; convolution.m:29-102):
 *   forward  y[i] = sum_{|d| <= M} k[d] x[i - d],
 *   adjoint  y[j] = sum_{|d| <= M} conj(k[d]) x[j + d],
 * with x = 0 outside the mix's N_f rows.
 * born_tgc_set: half_lengths holds M of every mix signal (measurement-major,
 * the order of the observations); kernels_c128 the concatenated 2 M + 1 taps
 * k[-M..M] of each, interleaved complex FP64.  A kernel with M + 1 > N_f is
 * cut symmetrically to 2 N_f - 1 taps, as convolution.m:76-90 does.
 * half_lengths = NULL clears it (regularization.tgc.off: identity).
 * born_tgc_apply: host FP64, in and out column-major N_obs x N_rhs
 * interleaved complex (in == out allowed), computed in FP64, synchronous.
 * born_tgc_apply_dev: complex64 device N_rhs x N_obs, in and out must not
 * overlap, enqueued on `stream`.  mode 1 = forward, any other value adjoint.
 * No counterpart in the MEX (a MATLAB-side linear transform). */
int born_tgc_set(born_plan_t plan, const int32_t* half_lengths, const double* kernels_c128);
int born_tgc_apply(born_plan_t plan, int32_t mode, const double* in, int32_t N_rhs, double* out);
int born_tgc_apply_dev(born_plan_t plan, int32_t mode, const void* in_c64, void* out_c64, int32_t N_rhs,
                       void* stream);

/* Received energies of one measurement under the plan's TGC (born_tgc_set):
 * energy_rx_canonical( op, LT_tgc ) with a diagonal LT_tgc
 * (operator_born.m:718-859, TGC at :836-843) -- for every grid point p, the
 * sum over the measurement's mixes of ||C_k Phi_k(:, p)||^2, Phi_k the mix's
 * rows of A at p and C_k its TGC convolution.  Host variant: N_points
 * doubles; device variant: N_points FP64 on the device, enqueued on `stream`
 * and complete on return.  BORN_ERR_INVALID_ARGUMENT when no TGC is set. */
int born_energy_rx_tgc(born_plan_t plan, int32_t measurement, double* energy);
int born_energy_rx_tgc_dev(born_plan_t plan, int32_t measurement, void* energy_f64, void* stream);

/* Copies of the device tables, complex64, for checking the device table
 * build: which = 0 h_ref (N_f_unique x N_h), 1 p_incident of measurement
 * `index` (N_f_unique(m) x N_points).  Column-major like the reference
 * matrices.  N_h = N_points for the reference's mirrored layout
 * (setup_grid_symmetric.m:95-241).  A plan built from geometry whose FOV is
 * narrower than the aperture (where that mirror map is undefined) keeps h_ref
 * on the grid extended laterally by (N_elements_axis - 1) pitches before the
 * FOV, indexed by lateral offset; then N_h is the size of that grid.
 * born_plan_table_shape returns the rows and columns of a table. */
int born_plan_table_shape(born_plan_t plan, int32_t which, int32_t index, int64_t* rows, int64_t* cols);
int born_plan_get_table(born_plan_t plan, int32_t which, int32_t index, void* out_c64, int64_t capacity);

/* Execution path chosen at plan creation: 0 structured (shift-invariant
 * lateral Toeplitz tiles), 1 general gather (any shift map). */
int born_plan_path(born_plan_t plan);

/* Force the general gather path (for cross-checking the structured one). */
int born_plan_force_general(born_plan_t plan, int32_t on);

/* Kernel timing with HIP events recorded on the launch stream around the
 * grid-contraction kernels (no counterpart in the reference; used by the
 * benchmark).  While enabled, every forward/adjoint apply records one event
 * pair per contraction launch.  born_plan_get_timing synchronizes and returns
 * the summed milliseconds and launch counts of [0] the forward contraction
 * and [1] the adjoint contraction, then clears them. */
int born_plan_set_timing(born_plan_t plan, int32_t on);
int born_plan_get_timing(born_plan_t plan, double* ms, int64_t* launches);

/* ---------------------------------------------------------------- batches
 * Several operator instances (plans of uneven FOV, frequency and wave
 * counts) applied together, one right-hand side each: a few persistent
 * contraction launches per direction, one per group of instances that take
 * the same kernel form, cover every instance.  No counterpart in
 * the reference (each MEX call serves one operator).  The plans must be on
 * the structured path, share the device, the precision, the element padding
 * and the points per pitch, and outlive the batch; at most 64 plans.  A plan
 * may still be applied on its own (with any rhs count): the next batch call
 * rebuilds the plan's one-rhs workspace and the batch's records.  A plan
 * whose precision was changed makes batch calls fail with born:Unsupported.
 * gamma_c64[i] / u_c64[i] are the device buffers of plan i (complex64,
 * N_points(i) and N_observations(i) entries). */
typedef struct born_batch_s* born_batch_t;
int born_batch_create(const born_plan_t* plans, int32_t n_plans, born_batch_t* batch);
int born_batch_destroy(born_batch_t batch);
int born_batch_forward_dev(born_batch_t batch, const void* const* gamma_c64, void* const* u_c64, void* stream);
int born_batch_adjoint_dev(born_batch_t batch, const void* const* u_c64, void* const* gamma_hat_c64, void* stream);
/* Replay through hipGraphs: the launch sequence of a direction is captured on
 * first use for a (stream, buffers) combination and replayed afterwards. */
int born_batch_set_graph(born_batch_t batch, int32_t on);
/* Autotuner: times the candidate work-item counts and persistent grid sizes
 * of the forward contraction and the segment widths of the adjoint on the
 * batch's own instances and keeps the fastest, for each MFMA launch set of
 * the batch; returns their times (ms, summed over the sets). */
int born_batch_autotune(born_batch_t batch, void* stream, double* fwd_ms, double* adj_ms);
int born_plan_autotune(born_plan_t plan, void* stream, double* fwd_ms, double* adj_ms);
/* Event timing of the batched contraction launches (as born_plan_set_timing). */
int born_batch_set_timing(born_batch_t batch, int32_t on);
int born_batch_get_timing(born_batch_t batch, double* ms, int64_t* launches);

/* Thread-local message of the last failure on this thread. */
const char* born_last_error(void);

/* Reference error identifier for a status code. */
const char* born_status_id(int status);

/* Error-path testing: the n-th device allocation from now fails with
 * BORN_ERR_OUT_OF_MEMORY (n = 0: the next one); n < 0 switches it off.
 * Process-wide.  No counterpart in the reference; it lets tests force the
 * allocation failures that the MEX reports as adjoint_quick_gpu:ErrorCUDA
 * (combined_quick_gpu.h:31-44) and check that every entry point cleans up. */
int born_debug_fail_allocation(int64_t n);
/* Number of device allocations the library holds now (all plans and
 * batches); with born_debug_fail_allocation it shows that a failed call
 * released what it had allocated.  No counterpart in the reference. */
int64_t born_debug_live_allocations(void);

#ifdef __cplusplus
}
#endif
#endif /* BORN_H */
