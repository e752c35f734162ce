/*
 * e2edp.h — C ABI of the MI355X-native end-to-end dynamic-programming planner.
 *
 * The library (libe2edp.so) replaces the two in-process host functions and the
 * two executables of the reference pipeline:
 *
 *   reference                                            replaced by
 *   ---------------------------------------------------  ---------------------------------
 *   build_sparse_transition_model_at_T_at_a(...)          e2e_model_build (all t, all a in
 *     src/build_model.cu:696-892 (+ the t x a host loop     one device-resident build)
 *     src/build_model.cu:1142-1165)
 *   save_master_Coos_to_file(...) src/build_model.cu:1246 e2e_model_export_coo[_i32]
 *   get_mdp_model(path)           src/spvi_2.cu:393-504   e2e_model_import_coo
 *   solver_spvi_solve(...)        src/spvi_2.cu:650-726   e2e_vi_solve
 *
 * Conventions
 *   - Every entry point returns an e2e_status; on failure e2e_last_error() gives a
 *     thread-local message. No C++ exception crosses this boundary.
 *   - Host buffers are owned by the caller. Device memory is owned by the library
 *     through opaque handles.
 *   - A handle is bound to one HIP stream (given at creation or created by the
 *     library). One handle must not be used by two host threads at once;
 *     distinct handles may.
 *   - All results are deterministic: no floating-point atomics touch results.
 */
#ifndef E2EDP_H
#define E2EDP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum e2e_status {
    E2E_OK = 0,
    E2E_ERR_INVALID = 1,      /* bad argument / inconsistent shapes */
    E2E_ERR_HIP = 2,          /* a HIP runtime call failed */
    E2E_ERR_NOMEM = 3,        /* device or host allocation failed */
    E2E_ERR_NOT_CONVERGED = 4,/* max_sweeps reached before the residual fell below thresh */
    E2E_ERR_TIMEOUT = 5,      /* wall-clock limit reached (spvi_2.cu:652 max_solver_time_s) */
    E2E_ERR_IO = 6,           /* file read/write or format error */
    E2E_ERR_CAPACITY = 7,     /* a size limit of this implementation was exceeded */
    E2E_ERR_FIELD = 8         /* non-finite velocity sample found in the input fields */
} e2e_status;

/* Problem parameters. Field-for-field the variables of src/input_to_build_model.h:3-48
 * plus alpha (src/build_model.cu:934-943) and the reward method code
 * (get_reward_type, src/build_model.cu:898-918). The reference packs them as a
 * 32-float vector (src/build_model.cu:1032-1055); e2e_params_pack() writes that
 * vector. */
typedef struct e2e_params {
    int32_t gsize;             /* grid cells per side (square grid)            h:11 */
    int32_t nt;                /* time steps                                    h:9  */
    int32_t nrzns;             /* Monte Carlo realizations                      h:23 */
    int32_t nmodes;            /* DO modes                                      h:41 */
    int32_t num_ac_speeds;     /*                                               h:19 */
    int32_t num_ac_angles;     /*                                               h:20 */
    int32_t i_term;            /* top-left row of the terminal subgrid          h:33 */
    int32_t j_term;            /* top-left column of the terminal subgrid       h:34 */
    int32_t term_subgrid_size; /* side of the terminal subgrid (0: none)        h:40 */
    int32_t reward_type;       /* 0 time,1 energy1,2 energy2,3 energy3,4 custom1     */
    int32_t neighb_gsize;      /* odd window side m (out-of-window is reported) h:44 */
    int32_t is_stationary;     /* carried into prob_params only                 h:46 */
    float F;                   /* max vehicle speed                             h:18 */
    float dt;                  /*                                               h:10 */
    float dx, dy;              /*                                               h:13 */
    float x0, y0;              /* centre of cell column 0 / row gsize-1         h:15 */
    float r_outbound;          /*                                               h:26 */
    float r_terminal;          /*                                               h:27 */
    float alpha;               /* energy/time mix of reward_type 4              */
    float z;                   /* filler of prob_params slots 20..31            h:48 */
} e2e_params;

/* Input fields, C-order, layouts of src/build_model.cu:380-398 and :978-985:
 *   u, v, s      : float [nt][gsize][gsize]
 *   ui, vi       : float [nt][nmodes][gsize][gsize]
 *   yi           : float [nt][nrzns][nmodes]
 *   mask         : int32 [nt][gsize][gsize]  (1 = obstacle)
 * s may be NULL when reward_type is 0, 1 or 4; mask may be NULL (no obstacles). */
typedef struct e2e_fields {
    const float* u;
    const float* v;
    const float* ui;
    const float* vi;
    const float* yi;
    const float* s;
    const int32_t* mask;
} e2e_fields;

typedef struct e2e_build_opts {
    int32_t fields_on_device;  /* 1: e2e_fields pointers are device pointers */
    int32_t keep_rsum;         /* 1: also keep per-row reward sums (needed by e2e_vi_solve_goals) */
    int32_t slot_k;            /* staging entries per row before the overflow arena (0: 16) */
    int32_t radix_bits;        /* digit width of the LDS segmented radix sort, 1..8 (0: 8) */
    int32_t slices_per_launch; /* time slices per build-kernel launch (0: all) */
    int32_t block_threads;     /* build workgroup size 256/512/1024 of both counting kernels (0: chosen from
                                  the LDS footprint and register use) */
    void* stream;              /* hipStream_t to run on; NULL: the library creates one */
    int32_t count_kernel;      /* 0: one wavefront per (state, action) row, rows it cannot take (S2 offsets
                                  beyond +-127 cells or more than 256 candidate keys) go to the block-level
                                  kernel; 1: the block-level kernel (one workgroup per row) for every row */
    int32_t paired_rounds;     /* rounds of 64 realizations in the per-row kernel: 0 automatic (paired when
                                  its LDS leaves fewer than 4 wavefronts per SIMD), 1 one at a time, 2 in
                                  pairs (the next round's LDS read issued ahead); other values are an error */
    int32_t n_alpha;           /* > 0: also build the rewards for these n_alpha values of alpha (at most 64;
                                  reward_type 4 only): the alpha sweep of src/my_script.sh:10-19 from one
                                  build, its transitions being the same for every alpha */
    int32_t lane_moves;        /* per-row kernel: 0 counts every cell the fast unit-cell pass can take at once
                                  (each realization placed among all actions' landing-cell boundaries,
                                  the same counts and rewards, bit for bit); 1 moves every (state, action,
                                  realization) lane; other values are an error */
    const float* alphas;       /* [n_alpha] host array, each alpha as the reference forms it
                                  (stof(argv[1]) / 100, src/build_model.cu:937-939) */
    int32_t comp_span;         /* the compressed entry stream the build writes beside the CSR entries (when
                                  nrzns < 65536) keeps a 256-entry window at 4 bytes per entry when its
                                  columns span less than this (0: 65536, at most 65536); lower values only
                                  serve to test windows kept at 8 bytes */
} e2e_build_opts;

typedef struct e2e_model_info {
    int64_t ns;                /* states = gsize^2 * nt                 (DP_relv_params[0]) */
    int64_t na;                /* actions = num_ac_speeds*num_ac_angles (DP_relv_params[1]) */
    int64_t nnz;               /* stored transitions                                        */
    int64_t nc;                /* cells per slice (0 for an imported model)                 */
    int64_t nt;                /* slices (0 for an imported model)                          */
    int64_t lanes;             /* Monte Carlo lanes evaluated = (nt-1)*nc*na*nrzns          */
    int64_t n_out_of_window;   /* lanes whose S2 fell outside the neighb_gsize window       */
    int64_t max_row_nnz;
    int64_t n_overflow_rows;   /* rows with more unique S2 than slot_k                     */
    int64_t n_multipass_rows;  /* rows whose radix sort needed more than one digit pass     */
    int64_t device_bytes;      /* device memory held by the model                          */
    double build_ms;           /* wall time of e2e_model_build                             */
    double move_kernel_ms;     /* summed HIP-event time of the move/sort/count kernel      */
    double scan_ms;            /* HIP-event time of the row-count scan                     */
    double csr_ms;             /* HIP-event time of writing the CSR and compressed entries */
    int32_t has_rsum;
    int32_t move_launches;
    int32_t block_threads;     /* workgroup size the build kernel ran with */
    int32_t unit_cells;        /* 1: the unit-cell (dx = dy = 1) move was used */
    int32_t wave_blocks;       /* wavefronts per block of the per-row kernel (0: not used) */
    int32_t n_deferred_rows;   /* rows the per-row kernel handed to the block-level kernel */
    int32_t paired_rounds;     /* 1: the per-row kernel ran its rounds in pairs */
    int32_t counted_cells;     /* (slice, cell) pairs whose rows came from one count over the cell's samples */
} e2e_model_info;

typedef struct e2e_solve_opts {
    float thresh;              /* stop at the first sweep whose sup-norm residual < thresh (spvi_2.cu:685) */
    float discount;            /* Q = R + discount * P V (spvi_2.cu:61,189) */
    int32_t max_sweeps;        /* 0: nt + 2 for a built model, 100000 for an imported one */
    int32_t sweeps_per_check;  /* sweeps launched between host checks of the converged flag (0: checks after
                                  16, 32, 64, 128 sweeps and then every 64; batched solves after 8, 16,
                                  32, 64 and then every 64) */
    int32_t use_graph;         /* 1: replay the sweep chain as a captured hipGraph */
    int32_t loads_in_flight;   /* nnz loads per lane per window: 2, 4 or 8 (0: chosen from nnz per state) */
    double timeout_ms;         /* 0: none */
    int32_t kernel_variant;    /* sweep kernel; 0: default (6 for a built model with nrzns < 65536, else 3).
                                  1 chunked + compensated float, 2 chunked + fma float (faster; exceeds 1e-5
                                  over long horizons), 3 lean streaming + float64, 4 tiled with an LDS value
                                  band + float64, 5 jagged-diagonal + float64, 6 lean streaming over 4-byte
                                  compressed entries + float64 (built models only), 7 backward pass (built
                                  models only): one launch per slice from the last to the first, the fixed
                                  point the sweeps reach after nt sweeps, bit for bit, in one pass (thresh
                                  and max_sweeps unused; stats.sweeps = 1); other values are an error */
    int32_t blocks_per_cu;     /* resident sweep blocks per compute unit (0: chosen by the library) */
    int32_t comp_span;         /* variant 6: a 256-entry window whose columns span at least this many keeps
                                  its 8-byte entries (0: 65536, the most a 16-bit offset reaches; 1..65536) */
    int32_t full_sweeps;       /* 1: every sweep recomputes every state. 0 (default): on a built model the
                                  lean sweeps (variants 3 and 6) recompute, at sweep k >= 3, only the slices
                                  t < nt - k + 2; the others already hold their final values (transitions go
                                  from each slice to the next), so values, policies, residuals and the sweep
                                  count are the same bits */
} e2e_solve_opts;

typedef struct e2e_solve_stats {
    int32_t sweeps;            /* sweeps performed (the reference's num_iterations) */
    int32_t converged;
    float final_residual;      /* residual of the last sweep */
    int32_t loads_in_flight;
    double solve_ms;           /* wall time of the solve */
    double sweep_kernel_ms;    /* HIP-event time of all sweep launches (incl. early-exit ones) */
    double sweep_kernel_ms_per_sweep; /* sweep_kernel_ms / launches that did work */
    double bytes_per_sweep;    /* algorithmic bytes of one sweep in the reference's CSR terms:
                                  8 nnz + 4 (rows+1) + 4 rows + 12 ns (entries, row pointers, R, V read,
                                  V write, policy write); for sweeps over the active slices, the mean over
                                  the sweeps of the same terms for the slices each one recomputed */
    double moved_bytes_per_sweep; /* the same traffic in the layout the variant reads: 4-byte compressed
                                  entries (8 for windows kept uncompressed) and float64 values */
    int32_t launches;
    int32_t kernel_variant;    /* variant used (1-based, as in e2e_solve_opts) */
    int32_t blocks;            /* sweep grid size in blocks */
    int32_t compressed_windows; /* 256-entry windows the compressed variant reads at 4 bytes per entry */
} e2e_solve_stats;

typedef struct e2e_batch_opts {
    int32_t variant;           /* SpMM kernel variant (0..37; 27..37 read the next slice's values from an
                                  LDS halo of a tile of cells, built models only; 33..37 with one lane per
                                  row over 8 or 16 columns); -1: chosen by the library (by the column
                                  count on models of 16M entries or more with up to 16 columns,
                                  else timed on this model). -3: every column by one exact backward
                                  pass (kernel_variant 7's schedule), ending at its fixed point. Alpha
                                  columns also take -2: one plain sweep chain per column over the
                                  shared model (and run -3 as one backward pass per column) */
    int32_t use_graph;         /* 1: replay the sweep chain as a captured hipGraph */
    int32_t reserved[2];
} e2e_batch_opts;

typedef struct e2e_batch_stats {
    int32_t sweeps;            /* sweeps run until the last column converged */
    int32_t converged;         /* columns that converged */
    int32_t variant;           /* kernel variant used */
    int32_t cols_per_lane;     /* its column width per lane */
    int32_t states_per_wave;   /* its states (= na rows each) per wavefront */
    int32_t col_block;         /* columns covered by one wavefront */
    double solve_ms;
    double sweep_kernel_ms;
    double tune_ms;            /* autotuning time included in solve_ms (0 when cached) */
    double bytes_per_sweep;    /* algorithmic bytes of one batched sweep */
    int32_t launches;
    int32_t reserved;
} e2e_batch_stats;

typedef struct e2e_server_opts {
    int64_t hbm_budget_bytes;  /* device bytes for models + batches; 0: 80% of free memory */
    int32_t max_cols;          /* goal columns per model batch, 1..64 (0: 16) */
    int32_t sweeps_per_step;   /* sweeps launched per scheduler step on the running model with the most
                                  entries (0: 4); a model with k times fewer entries gets k times as many,
                                  up to 16 times */
    int32_t keep_full_results; /* 1: keep each request's full value/policy columns on the host */
    int32_t serial;            /* 1: one request on the GPU at a time (queue-only baseline) */
    float thresh;              /* VI stopping threshold (0: 1e-6) */
    int32_t exact_pass;        /* 1: each scheduler step runs one exact backward pass over a model's
                                  running columns (one launch per slice, last to first), which ends
                                  every one of them at its fixed point; 0: sweeps to thresh */
    int32_t max_models;        /* models whose steps run at once: the ones holding the oldest running
                                  requests (0: 1; -1: every model with running requests) */
} e2e_server_opts;

typedef struct e2e_request_result {
    int32_t status;            /* 0 queued, 1 running, 2 done, 3 rejected (too big), 4 failed */
    int32_t sweeps;            /* sweeps of this request's solve */
    float value_at_start;      /* V(start cell, t = 0) */
    int32_t action_at_start;   /* pi(start cell, t = 0) */
    double latency_ms;         /* arrival -> result */
    double queue_ms;           /* arrival -> admitted into a batch slot */
    int32_t cache_hit;         /* its model was resident when it arrived */
    int32_t reserved;
} e2e_request_result;

typedef struct e2e_server_stats {
    int64_t submitted, completed, rejected, models_built, model_evictions, cache_hits;
    int64_t peak_reserved_bytes, sweeps_launched;
    double p50_ms, p99_ms, mean_ms, max_ms;
    double wall_ms;            /* first arrival -> last completion */
    double solves_per_s;
} e2e_server_stats;

const char* e2e_last_error(void);
const char* e2e_version(void);

/* Number of HIP devices visible (0 without a GPU; never fails). */
int32_t e2e_device_count(void);

/* Device memory the library keeps mapped after models close (no reference counterpart:
 * the reference allocates and frees per executable run). Freed blocks of 1 MiB or more
 * stay mapped, up to E2E_POOL_BYTES bytes (environment; default 32 GiB, 0 = off), for
 * later builds and solves of similar size. e2e_cached_device_bytes reports the idle
 * bytes; e2e_release_cached_memory returns them to the HIP runtime. */
int64_t e2e_cached_device_bytes(void);
void e2e_release_cached_memory(void);

/* sizeof of the ABI structs, for binding checks: 0 e2e_params, 1 e2e_fields,
 * 2 e2e_build_opts, 3 e2e_model_info, 4 e2e_solve_opts, 5 e2e_solve_stats,
 * 6 e2e_batch_opts, 7 e2e_batch_stats, 8 e2e_server_opts, 9 e2e_request_result,
 * 10 e2e_server_stats;
 * -1 for an unknown code. */
int32_t e2e_struct_size(int32_t which);

/* Writes the reference's 32-float prob_params vector (src/build_model.cu:1032-1055). */
e2e_status e2e_params_pack(const e2e_params* p, float out32[32]);

/* Action table of src/build_model.cu:1403-1457: action = speed_idx*num_ac_angles + angle_idx.
 * speed/angle may be NULL. */
e2e_status e2e_actions(const e2e_params* p, float* speed, float* angle);

/* Monte Carlo build of the sparse transition model on the GPU.
 * Replaces the per-(t,a) build_sparse_transition_model_at_T_at_a
 * (src/build_model.cu:696-892) and the t x a loop of main (:1142-1165). */
e2e_status e2e_model_build(const e2e_params* p, const e2e_fields* f,
                           const e2e_build_opts* opts, void** model_out);

/* Loads a model given in the reference COO layout (rows a*Ns+s, sorted by row,
 * S2 ascending within a row) — what get_mdp_model reads (src/spvi_2.cu:393-504).
 * Host arrays. */
e2e_status e2e_model_import_coo(int64_t ns, int64_t na, int64_t nnz,
                                const int32_t* S1, const int32_t* S2,
                                const float* val, const float* R,
                                void* stream, void** model_out);

e2e_status e2e_model_info_get(const void* model, e2e_model_info* out);
e2e_status e2e_model_destroy(void* model);

/* Writes the model in the reference's on-disk order (src/build_model.cu:1246-1314):
 * rows a*Ns + t*g^2 + sp sorted by (a,t,sp), S2 ascending; R of length Na*Ns with
 * the last slice 0. Host buffers of nnz (S1,S2,val) and na*ns (R) elements.
 * Any pointer may be NULL to skip it. */
e2e_status e2e_model_export_coo(const void* model, int64_t* S1, int64_t* S2,
                                float* val, float* R);
/* Same, int32 indices (src/convert_model_datatype.py:19-23). */
e2e_status e2e_model_export_coo_i32(const void* model, int32_t* S1, int32_t* S2,
                                    float* val, float* R);

/* Internal state-major CSR (row = s*Na + a): row_ptr[na*ns+1], col[nnz], val[nnz],
 * R[na*ns]. Host buffers; any may be NULL. */
e2e_status e2e_model_export_csr(const void* model, uint32_t* row_ptr, int32_t* col,
                                float* val, float* R);

/* The reward vector of alpha column j of a model built with n_alpha > 0, in the
 * reference's order (R of length Na*Ns, rows a*Ns + s, the last slice 0): the
 * master_R that src/build_model.cu:1246-1314 writes for a build with that alpha
 * (src/my_script.sh:10-19 runs one build per alpha; its transitions do not depend on
 * alpha). Host buffer of na*ns floats. */
e2e_status e2e_model_export_R_alpha(const void* model, int32_t j, float* R);

/* Sparse value iteration: replaces solver_spvi_solve (src/spvi_2.cu:650-726).
 * policy/value are host arrays of ns elements (either may be NULL). */
e2e_status e2e_vi_solve(void* model, const e2e_solve_opts* opts,
                        uint32_t* policy, float* value, e2e_solve_stats* stats);

/* Sparse value iteration without device allocation. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results as e2e_vi_solve, bit for bit, for kernel_variant 0, 3 or 6 (6 needs the
 * compressed entries a build makes when nrzns < 65536; 0 takes them when there).
 * It calls no hipMalloc, hipFree or other allocating routine: every device buffer beyond
 * the model and the host outputs sits in one fixed array per GPU (56 MiB, held
 * until the process ends). Limits: ns <= 2,097,152 states; for a built model swept over
 * its active slices (full_sweeps 0), nt <= 256; use_graph must be 0. Larger problems
 * return E2E_ERR_CAPACITY. The stats carry no kernel times or byte counts (0). Calls
 * that overlap on one GPU share its array and race; run them one at a time per GPU. */
e2e_status e2e_vi_solve_noalloc(void* model, const e2e_solve_opts* opts,
                                uint32_t* policy, float* value, e2e_solve_stats* stats);

/* Batched multi-goal solve (no counterpart call in the reference, which rebuilds and
 * re-solves the model per goal: src/input_to_build_model.h:33-40 fix one terminal per
 * build). Column j equals the solve of the model built with its terminal cell at
 * goal_cells[j] (cell index i*gsize + j, term_subgrid_size 1), by solver_spvi_solve
 * (src/spvi_2.cu:650-726). The model must come from e2e_model_build with keep_rsum=1 and
 * term_subgrid_size 0, without obstacles. k <= 64. policy/value are host arrays [k][ns]
 * (either may be NULL); col_sweeps [k] gets each column's sweep count (or NULL). */
e2e_status e2e_vi_solve_goals(void* model, const int32_t* goal_cells, int32_t k,
                              const e2e_solve_opts* opts, const e2e_batch_opts* bopts,
                              uint32_t* policy, float* value, int32_t* col_sweeps,
                              e2e_batch_stats* stats);

/* Alpha-sweep solve: replaces the per-alpha spvi runs of src/my_script.sh:10-19
 * (src/Makefile:7-19, solver_spvi_solve src/spvi_2.cu:650-726 once per alpha). Column j
 * is the solve of the model built with alpha = alphas[j]; all columns share the one set
 * of transitions. The model must come from e2e_model_build with n_alpha > 0.
 * policy/value are host arrays [n_alpha][ns] (either may be NULL); col_sweeps
 * [n_alpha] gets each column's sweep count (or NULL). */
e2e_status e2e_vi_solve_alphas(void* model, const e2e_solve_opts* opts, const e2e_batch_opts* bopts,
                               uint32_t* policy, float* value, int32_t* col_sweeps,
                               e2e_batch_stats* stats);

/* Policy rollouts: the post-processing's trajectory simulation and metrics
 * (Post_processing/plot_e2eGPU_policy.py:376-520 plot_exact_trajectory_set_DP, its
 * vehicle step Post_processing/grid_world.py:160-230 Grid.move_exact, get_metrics
 * :776-797), one GPU lane per trajectory. */
typedef struct e2e_rollout_opts {
    int32_t interp;            /* 1: actions after the first bilinearly interpolated from the four
                                  nearest cells' actions (get_interpolated_action :174-198; the
                                  run the reference logs, :949-952); 0: the cell's own action */
    int32_t fields_on_device;  /* 1: e2e_fields pointers are device pointers */
    double cf, cr;             /* energy coefficients Cf, Cr (:427-428); 0: 1 */
    void* stream;              /* hipStream_t to run on; NULL: a stream of the call's own */
    int32_t reserved[4];
} e2e_rollout_opts;

typedef struct e2e_rollout_stats {
    int64_t trajectories;      /* n_starts x nrzns */
    int64_t reached;           /* trajectories that ended in the terminal subgrid */
    int64_t rejected;          /* interpolations at a support pattern the reference rejects
                                  (:158-161); the call then returns E2E_ERR_INVALID */
    int64_t moves;             /* vehicle steps over all trajectories */
    double kernel_ms;          /* the rollout kernel (HIP events on its stream) */
} e2e_rollout_stats;

/* Replaces the realization loop of plot_exact_trajectory_set_DP (:436-516) for any number
 * of start cells. policy: host int32 [nt][gsize][gsize] (policy.npy; -1 takes the last
 * action, as a Python list index does). starts: host int32 [n_starts][2] = (row i,
 * column j). Trajectory n = k*nrzns + r starts at start k under realization r. Outputs
 * (host, caller-owned): success [n] (1: reached the terminal subgrid, 0: left the grid
 * interior, entered an obstacle or ran out of time, -1: rejected interpolation), travel
 * [n] (moves), energy [n][3] (energy_cons, energy_col, net_energy_cons), xy
 * [n][nt+1][2] positions per step or NULL. opts may be NULL (interp 1). */
e2e_status e2e_rollout(const e2e_params* p, const e2e_fields* fields, const int32_t* policy,
                       int32_t n_starts, const int32_t* starts, const e2e_rollout_opts* opts,
                       int8_t* success, int32_t* travel, double* energy, double* xy,
                       e2e_rollout_stats* stats);

/* Input generation: the double-gyre environment of
 * src/data_input/build_environment_DG_with_westward_clouds.py (DG_velocity_field
 * :105-240, DG_scalar_field :20-100, fill_obstacles :470-505) at refinement degree 1:
 * realizations for n_wsamples forcing frequencies, the per-slice mean, the n_modes
 * leading modes and coefficients of the centred realization matrix (the truncated
 * np.linalg.svd of :227-236, computed as the leading eigenvectors of its Gram matrix),
 * the cloud field and the obstacle mask, written in the layouts e2e_fields reads. */
typedef struct e2e_env_params {
    int32_t gsize;             /* cells per side (init_gsize, interpolate_degree 1) */
    int32_t nt;                /* time slices */
    int32_t n_wsamples;        /* realizations: forcing frequencies across [w_lo, w_hi] (>= 2) */
    int32_t n_modes;           /* modes kept, 1..min(n_wsamples, gsize^2) */
    double dt, dxy;            /* slice time step, cell size */
    double A, eps, wx, wy;     /* double-gyre amplitude, perturbation, wavenumbers (the reference:
                                  wx given, wy = pi/2) */
    double w_lo, w_hi;         /* w_range */
    double A_sc;               /* cloud amplitude */
    int32_t obstacle_mode;     /* 0 none, 1 static box, 2 moving box */
    int32_t obs_upper_row, obs_left_col, obs_width;  /* mode 2: the box at t = 0 */
    double obs_speed_x, obs_speed_y;                  /* mode 2: cells per time step */
    int32_t reserved[4];
} e2e_env_params;

typedef struct e2e_env_out {
    float* u;                  /* [nt][g][g] all_u_mat (mean vx) */
    float* v;                  /* [nt][g][g] all_v_mat */
    float* ui;                 /* [nt][n_modes][g][g] all_ui_mat */
    float* vi;                 /* [nt][n_modes][g][g] all_vi_mat */
    float* yi;                 /* [nt][n_wsamples][n_modes] all_Yi */
    float* s;                  /* [nt][g][g] all_s_mat (clouds), or NULL */
    int32_t* mask;             /* [nt][g][g] obstacle_mask, or NULL */
    float* phi;                /* [nt][g][g] all_phi_mat; phi, phi_i, phi_yi all or none */
    float* phi_i;              /* [nt][n_modes][g][g] all_phi_i_mat */
    float* phi_yi;             /* [nt][n_wsamples][n_modes] all_phi_Yi_mat */
    double* singular_values;   /* [nt][n_modes] of the velocity matrix, or NULL */
} e2e_env_out;

typedef struct e2e_env_stats {
    double total_ms;
    double modes_ms;           /* Gram matrices, subspace iteration and mode assembly */
    int32_t slices_per_pass;
    int32_t max_iterations;    /* subspace iterations of the slowest pass */
} e2e_env_stats;

/* Host output arrays (caller-owned). Modes and coefficients carry the sign that makes
 * each mode's first largest-magnitude entry positive (LAPACK's sign is arbitrary;
 * every coefficient x mode product is unchanged by it). */
e2e_status e2e_env_generate(const e2e_env_params* p, const e2e_env_out* out, e2e_env_stats* stats);

/* Continuous-batching planning server (no counterpart in the reference, whose pipeline
 * is one build + one solve per goal: src/Makefile:7-19). Environments (fields + params,
 * keyed by env_id, e.g. the field seed) are registered once; requests (env, goal cell,
 * start cell at t = 0) are answered from device-resident models. */
e2e_status e2e_server_create(const e2e_server_opts* opts, void** server_out);
e2e_status e2e_server_add_env(void* server, int64_t env_id, const e2e_params* p,
                              const e2e_fields* host_fields);
e2e_status e2e_server_submit(void* server, int64_t env_id, int32_t goal_cell,
                             int32_t start_cell, uint64_t* req_id);
e2e_status e2e_server_poll(void* server, uint64_t req_id, e2e_request_result* out);
e2e_status e2e_server_wait(void* server, uint64_t req_id, double timeout_ms,
                           e2e_request_result* out);
/* Full value/policy columns (ns entries) of a finished request (keep_full_results=1). */
e2e_status e2e_server_result_columns(void* server, uint64_t req_id, uint32_t* policy,
                                     float* value);
e2e_status e2e_server_drain(void* server, double timeout_ms);
e2e_status e2e_server_stats_get(void* server, e2e_server_stats* out);
e2e_status e2e_server_destroy(void* server);

#ifdef __cplusplus
}
#endif

#endif /* E2EDP_H */
