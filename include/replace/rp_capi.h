/*
 * rp_capi.h -- C ABI of the MI355X (gfx950) Nesterov global placer.
 *
 * This is the drop-in boundary underneath the replace::Replace C++ facade
 * (include/replace/Replace.h).  Every entry point takes plain pointers and
 * sizes, never C++ or torch types, so any host language can bind it
 * (see INTEGRATION.md for the ctypes / C++ bindings).
 *
 * Reference interface being replaced (reference = The-OpenROAD-Project/RePlAce):
 *   rp_create           <- PlacerBase::PlacerBase + PlacerBase::init   src/placerBase.cpp:592-758
 *                          (OpenDB traversal replaced by the host SoA rp_netlist below)
 *   rp_initial_place    <- Replace::doInitialPlace                     src/replace.cpp:73-88
 *                          InitialPlace::doBicgstabPlace               src/initialPlace.cpp:56-105
 *   rp_nesterov_place   <- Replace::doNesterovPlace                    src/replace.cpp:90-128
 *                          NesterovBase::init, NesterovPlace::init     src/nesterovBase.cpp:916-1062,
 *                                                                      src/nesterovPlace.cpp:64-175
 *                          NesterovPlace::doNesterovPlace              src/nesterovPlace.cpp:275-472
 *   rp_get_positions    <- NesterovPlace::updateDb / dbInst::setLocation src/nesterovPlace.cpp:585-595
 *   rp_set_positions    <- incremental mode: positions already in the DB  src/initialPlace.cpp:67-69
 *   rp_destroy          <- Replace::~Replace / reset                   src/replace.cpp:35-65
 *
 * Error convention.  The reference prints "[ERROR] msg (REPL-code)" and
 * calls exit(code) (src/logger.cpp:31-36).  This ABI never exits: it returns
 *   0            success
 *   1..4         the reference's REPL error codes with the same meaning
 *                (1 negative filler area / utilisation > 100.1 %,
 *                 2 diverged during Nesterov init, 3 back-tracking exhausted,
 *                 4 overflow rebound with HPWL growth)
 *   < 0          HIP / resource / argument errors; rp_last_error() has text.
 * The facade maps 1..4 back to the reference message and exit code.
 */
#ifndef REPLACE_RP_CAPI_H
#define REPLACE_RP_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RP_OK 0
#define RP_ERR_FILLER 1      /* REPL-1 */
#define RP_ERR_INIT_DIVERGED 2 /* REPL-2 */
#define RP_ERR_BACKTRACK 3   /* REPL-3 */
#define RP_ERR_REBOUND 4     /* REPL-4 */
#define RP_ERR_ARG (-1)
#define RP_ERR_HIP (-2)
#define RP_ERR_NOMEM (-3)
#define RP_ERR_STATE (-4)
#define RP_ERR_NO_DEVICE (-5)

/*
 * Host SoA netlist.  Borrowed for the duration of rp_create only; the
 * library copies what it needs.  Units are database units (DBU).
 *
 * Mirrors what PlacerBase::init reads from OpenDB (src/placerBase.cpp:605-758):
 *  - rows: the core area is the bounding box of all rows
 *    (getCoreRectFromDb, src/placerBase.cpp:961-976) and rows drive the
 *    fragmented-row dummy instances (src/placerBase.cpp:760-846).
 *  - insts: every dbInst in block order.  inst_box is the placement bbox
 *    (location + master size).  inst_fixed = LOCKED/FIRM/COVER status.
 *  - nets / pins: signal nets only (POWER/GROUND/RESET skipped by the
 *    caller, src/placerBase.cpp:686-693), pins net-major: a net's ITerms
 *    first, then its BTerms (src/placerBase.cpp:702-713).
 *  - pin_inst[p] >= 0: ITerm of that instance; pin_off[2p..2p+1] is the
 *    offset of the pin-shape bbox centre from the master centre
 *    (Pin::updateCoordi(dbITerm*), src/placerBase.cpp:330-376).
 *  - pin_inst[p] == -1: top-level port (BTerm); pin_box[4p..4p+3] is the
 *    bbox of its BPins, or (INT_MAX, INT_MAX, INT_MIN, INT_MIN) when the
 *    port is unplaced (src/placerBase.cpp:381-406).  pin_box may be NULL
 *    when the netlist has no BTerms.
 *  - inst_pin_start/inst_pins (optional, NULL = net order): per-instance
 *    ITerm order (dbInst::getITerms, src/placerBase.cpp:729-743).
 */
typedef struct rp_netlist {
  int32_t dbu_per_micron;
  int32_t die_lx, die_ly, die_ux, die_uy;
  int32_t site_w, site_h;
  int32_t num_rows;
  const int32_t* row_box;       /* 4 * num_rows */
  int32_t num_insts;
  const int32_t* inst_box;      /* 4 * num_insts */
  const uint8_t* inst_fixed;    /* num_insts */
  int32_t num_nets;
  const int32_t* net_pin_start; /* num_nets + 1 */
  int32_t num_pins;
  const int32_t* pin_inst;      /* num_pins */
  const int32_t* pin_off;       /* 2 * num_pins */
  const int32_t* pin_box;       /* 4 * num_pins or NULL */
  const int32_t* inst_pin_start;/* num_insts + 1 or NULL */
  const int32_t* inst_pins;     /* inst_pin_start[num_insts] or NULL */
} rp_netlist;

/*
 * Every Replace setter (include/replace/Replace.h:39-61) plus the constants
 * the reference keeps without a setter (src/nesterovPlace.cpp:19-29,
 * src/nesterovBase.cpp:875-880).  rp_params_default() gives the values of
 * the Replace constructor (src/replace.cpp:13-33).
 */
typedef struct rp_params {
  int32_t initial_place_max_iter;         /* 20   */
  int32_t initial_place_min_diff_length;  /* 1500 */
  int32_t initial_place_max_solver_iter;  /* 100  */
  int32_t initial_place_max_fanout;       /* 200  */
  float initial_place_net_weight_scale;   /* 800  */
  int32_t nesterov_place_max_iter;        /* 2000 */
  int32_t bin_grid_cnt_x;                 /* 0 = automatic */
  int32_t bin_grid_cnt_y;                 /* 0 = automatic */
  float target_overflow;                  /* 0.1  */
  float target_density;                   /* 1.0  */
  float init_density_penalty;             /* 8e-5 */
  float init_wire_length_coef;            /* 0.25 */
  float min_phi_coef;                     /* 0.95 */
  float max_phi_coef;                     /* 1.05 */
  float reference_hpwl;                   /* 446000000 */
  int32_t incremental_place_mode;         /* 0 */
  int32_t verbose;                        /* 0 */
  /* constants without a reference setter */
  int32_t max_back_track;                 /* 10 */
  float min_preconditioner;               /* 1.0 */
  float initial_prev_coordi_update_coef;  /* 100 */
  float min_wire_length_force_bar;        /* -300 */
  /* Poisson transform (no reference counterpart; additive):
   *   0  FP32 FFT-based DCT/DST (default; the reference's fftsg arithmetic class)
   *   1  dense transform-matrix GEMMs on matrix cores, split-BF16 operands
   *      (hi + lo) with FP32 accumulation, run by hipBLASLt */
  int32_t poisson_transform;              /* 0 */
} rp_params;

/* Per-iteration record of the Nesterov loop: the values the reference prints
 * in "[NesterovSolve] Iter: i overflow: o HPWL: h" (src/nesterovPlace.cpp:421-423). */
typedef struct rp_iter_record {
  float overflow;
  float step_length;
  float density_penalty;
  int32_t num_back_track;
  int64_t hpwl;
} rp_iter_record;

typedef struct rp_nesterov_stats {
  int32_t iterations;     /* outer iterations run (the i+1 of the last one) */
  int32_t evals;          /* gradient evaluations including the two init evals */
  int32_t finished;       /* 1 when the overflow target was reached */
  int32_t code;           /* 0 or REPL code 2/3/4 */
  float final_overflow;
  float final_step_length;
  float final_density_penalty;
  int32_t diverged;       /* 1 when the loop stopped on divergence: code 2, 3, 4,
                             or 0 for a NaN / inf gradient sum in the loop
                             (the reference's "[ERROR]  (REPL-0)" exit,
                             src/nesterovPlace.cpp:371-374, 466-471) */
  int64_t final_hpwl;     /* pin-centre HPWL of the last accepted iterate */
  double seconds;         /* device time of the loop, init included */
} rp_nesterov_stats;

typedef struct rp_ip_stats {
  int32_t iterations;
  float last_error;
  int64_t last_hpwl;     /* inst-bbox HPWL printed at the last iteration */
  double seconds;
} rp_ip_stats;

typedef struct rp_ctx rp_ctx;

/* The values PlacerBase::init and PlacerBase::printInfo log
 * (src/placerBase.cpp:609-626, 902-958): "[INFO] DBU", "SiteSize",
 * "CoreArea*", "NumInstances" ... "MacroInstsArea".  num_insts counts the
 * fragmented-row dummy instances; num_fixed_insts the fixed instances that
 * overlap the core. */
typedef struct rp_pb_info {
  int32_t dbu_per_micron, site_w, site_h;
  int32_t die_lx, die_ly, die_ux, die_uy;
  int32_t core_lx, core_ly, core_ux, core_uy;
  int32_t num_insts, num_place_insts, num_fixed_insts, num_dummy_insts, num_nets, num_pins;
  int32_t num_unplaced_ports;   /* BTerm pins without a placed shape (the "[WARN] ... not placed" lines) */
  int64_t core_area, nonplace_insts_area, place_insts_area, std_insts_area, macro_insts_area;
  float util;                   /* percent, as "Util(%)" prints it */
} rp_pb_info;

void rp_params_default(rp_params* params);

/* PlacerBase::init: instance classification, areas, fragmented-row dummy
 * instances, pins.  Returns 1 (REPL exit 1) when utilisation >= 100.1 %.
 * Device selection: the current HIP device. */
int rp_create(const rp_netlist* netlist, const rp_params* params, rp_ctx** out);

/* PlacerBase values of a context (valid after rp_create, also when it
 * returned REPL-1). */
int rp_get_pb_info(rp_ctx* ctx, rp_pb_info* info);

/* Replace::doInitialPlace (B2B net model + BiCGSTAB on the device). */
int rp_initial_place(rp_ctx* ctx, rp_ip_stats* stats);

/* Replace::doNesterovPlace: NesterovBase + NesterovPlace init and the loop.
 * Returns 0 or REPL code 1..4; positions are updated in every case, as the
 * reference calls updateDb before exiting (src/nesterovPlace.cpp:465-471). */
int rp_nesterov_place(rp_ctx* ctx, rp_nesterov_stats* stats);

/* ------------------------------------------------------------------------
 * The Nesterov loop in pieces: init, a bounded number of outer iterations,
 * and the loop state between two iterations.  rp_nesterov_place is
 * rp_nesterov_init followed by rp_nesterov_iterate(max_iter) (with its
 * positions written back).  Getting and setting the state lets a run be
 * checkpointed and resumed, and a single iteration be checked from a given
 * state (teacher forcing against the oracle).
 * ---------------------------------------------------------------------- */

/* NesterovPlace's state between two outer iterations (the members of
 * src/nesterovPlace.h:76-127 the loop reads, and the loop locals of
 * doNesterovPlace, src/nesterovPlace.cpp:300-313), per-gcell arrays in the
 * reference's gcell order with x, y interleaved (2 * num_gcells floats).
 * Array pointers may be NULL (not read / not written). */
typedef struct rp_nesterov_state {
  float* cur_slp;          /* curSLPCoordi_ */
  float* cur_slp_grad;     /* curSLPSumGrads_ (preconditioned sum gradient) */
  float* prev_slp;         /* prevSLPCoordi_ */
  float* prev_slp_grad;    /* prevSLPSumGrads_ */
  float* cur_coord;        /* curCoordi_ */
  float step_length;       /* stepLength_ */
  float density_penalty;   /* densityPenalty_ */
  float wl_coef_x, wl_coef_y;  /* wireLengthCoefX_/Y_ */
  float base_wl_coef;      /* baseWireLengthCoef_ */
  float cur_a;             /* curA: a_k of the next iteration's prevA */
  float sum_overflow;      /* sumOverflow_ */
  float min_sum_overflow;  /* loop local minSumOverflow */
  float hpwl_with_min_sum_overflow;  /* loop local hpwlWithMinSumOverflow */
  float max_phi_coef;      /* npVars_.maxPhiCoef (scaled once by 0.99) */
  int32_t is_max_phi_coef_changed;   /* loop local */
  int32_t iter;            /* outer iterations done (the next one is iter + 1) */
  int64_t prev_hpwl;       /* prevHpwl_ */
} rp_nesterov_state;

/* NesterovPlace::init (src/nesterovPlace.cpp:64-175): the two lambda-0
 * evaluations, initial penalty and step length.  Returns 0, 2 (diverged in
 * init) or < 0. */
int rp_nesterov_init(rp_ctx* ctx);
/* Up to n_iters more outer iterations of doNesterovPlace
 * (src/nesterovPlace.cpp:315-463) from the current state, stopping early on
 * the overflow target, divergence or nesterov_place_max_iter; positions are
 * updated (updateDb).  stats covers the run since rp_nesterov_init. */
int rp_nesterov_iterate(rp_ctx* ctx, int32_t n_iters, rp_nesterov_stats* stats);
int rp_nesterov_get_state(rp_ctx* ctx, rp_nesterov_state* state);
/* Replaces the loop state (valid after rp_nesterov_init); the next
 * rp_nesterov_iterate starts iteration state->iter + 1 from it. */
int rp_nesterov_set_state(rp_ctx* ctx, const rp_nesterov_state* state);

/* [InitialPlace] lines of the last rp_initial_place call: CG error and the
 * inst-bbox HPWL printed per iteration (src/initialPlace.cpp:88-90). */
int rp_get_ip_records(rp_ctx* ctx, float* err, int64_t* hpwl, int32_t capacity, int32_t* count);

/* Per-iteration records of the last rp_nesterov_place call. */
int rp_get_iter_records(rp_ctx* ctx, rp_iter_record* out, int32_t capacity, int32_t* count);

/* Lower-left corners of all num_insts instances (as the DB would hold them). */
int rp_get_positions(rp_ctx* ctx, int32_t* lx, int32_t* ly);

/* Density centres (GCell dCx, dCy: the integer centre of the last evaluated
 * iterate, src/nesterovBase.cpp:164-179) of every gcell, in the reference's
 * gcell order (placed instances, then fillers; rp_nb_info.num_gcells
 * entries each).  No reference counterpart as an API: it is the state the
 * reference's PlotEnv draws (src/plot.cpp:284-426), for debug dumps. */
int rp_get_gcell_centres(rp_ctx* ctx, float* cx, float* cy);
/* Move instances (incremental mode / resume); lx, ly for all num_insts. */
int rp_set_positions(rp_ctx* ctx, const int32_t* lx, const int32_t* ly);

void rp_destroy(rp_ctx* ctx);
const char* rp_last_error(rp_ctx* ctx);

/* Verbose-3 trace (rp_params.verbose >= 3): one record per gradient
 * evaluation of the last rp_nesterov_place / rp_nesterov_iterate run, with
 * the values the reference logs at verbose level 3 in updateGradients,
 * getStepLength and updateNextIter (src/nesterovPlace.cpp:206-273, 514-571):
 * density penalty, the three gradient sums, the two distances and the new
 * step length, the new gradient's second norm, sum phi * area (sumPhi,
 * src/nesterovBase.cpp:1493-1505; 0 with the matrix-core transform), the
 * overflow and the HPWL of the evaluated iterate.  iter is the outer
 * iteration (0-based), -1 for the two NesterovPlace::init evaluations;
 * trial the back-tracking trial.  Sums are FP64 reductions, so values
 * differ from the reference's float sums in the last digits. */
typedef struct rp_eval_record {
  int32_t iter, trial, diverged;
  float density_penalty, wl_coef;
  float wl_grad_sum, density_grad_sum, grad_sum;
  float coordinate_distance, gradient_distance, new_step_length;
  float gradient_norm, sum_phi, overflow;
  int64_t hpwl;
} rp_eval_record;
int rp_get_eval_records(rp_ctx* ctx, rp_eval_record* out, int32_t capacity, int32_t* count);

/* Per-kernel timing of the evaluation slots (no reference counterpart: the
 * reference has no timers).  When enabled, rp_nesterov_place brackets every
 * kernel of every executed slot with HIP events on the context's stream and
 * accumulates the elapsed time (enabling clears the totals, disabling keeps
 * them; the events themselves slow the run, so time whole placements with
 * timing off).  rp_get_kernel_timing fills 11 entries:
 * for the 10 kernel slots in launch order (step, density scatter, a
 * reserved slot (always 0: the WA kernel computes the pin centres),
 * bin finalize + row DCT, matrix-core Poisson solve (0 with the FFT
 * transform), column spectral pass, row inverse (both 0 with the
 * matrix-core transform), WA, gradient, controller) the summed
 * milliseconds, then the summed milliseconds of the spatial-order refresh
 * that runs once per batch of queued slots; *launches is the number of
 * executed slots. */
int rp_set_kernel_timing(rp_ctx* ctx, int enable);
int rp_get_kernel_timing(rp_ctx* ctx, double* total_ms_per_kernel, int64_t* launches);

/* ------------------------------------------------------------------------
 * Stage-level entry points.  They expose single pieces of one gradient
 * evaluation (src/nesterovPlace.cpp:338-377) so parity can be checked one
 * stage at a time against the same state.  Valid after rp_nesterov_setup.
 * ---------------------------------------------------------------------- */

/* Geometry of the Nesterov state built by NesterovBase::init. */
typedef struct rp_nb_info {
  int32_t num_gcells, num_fillers, num_place_insts;
  int32_t num_gnets, num_gpins;
  int32_t bin_cnt_x, bin_cnt_y, bin_size_x, bin_size_y;
  int32_t core_lx, core_ly, core_ux, core_uy;
  int32_t filler_dx, filler_dy;
  int64_t std_insts_area, macro_insts_area, place_insts_area, nonplace_insts_area;
  float target_density;
  /* the values NesterovBase::init logs (src/nesterovBase.cpp:645-649,
   * 1127-1133): FillerInit areas and the automatic bin-count inputs */
  int64_t core_area, white_space_area, movable_area, total_filler_area;
  int64_t average_place_inst_area, ideal_bin_area, ideal_bin_count, total_bin_area;
} rp_nb_info;

/* NesterovBase::init only (gcells, fillers, bin grid, density sizes). */
int rp_nesterov_setup(rp_ctx* ctx, rp_nb_info* info);

/* Per-gcell constants: density box half sizes, density scale, cell size,
 * initial density centre (after updateDensityCoordiLayoutInside). */
int rp_get_gcells(rp_ctx* ctx, int32_t* dhalf_x, int32_t* dhalf_y, float* scale,
                  int32_t* dx, int32_t* dy, float* init_cx, float* init_cy, uint8_t* kind);

/* Bin nonPlace areas (int64, B). */
int rp_get_nonplace(rp_ctx* ctx, int64_t* nonplace);

/* One gradient evaluation at density centres (cx, cy) (float, per gcell),
 * density penalty lambda and wire-length coefficients.  Any output pointer
 * may be NULL.  Outputs are host buffers:
 *   bin_inst, bin_filler : int64[B]   scattered areas (bins y-major, j*Nx+i)
 *   density, phi, ex, ey : float[B]
 *   overflow_area        : int64[1]
 *   net_sums             : float[8*Nn] (minX, xminX, maxX, xmaxX, minY, yminY, maxY, ymaxY)
 *   wl_grad, dens_grad, sum_grad : float[2*Nc] (x, y interleaved)
 *   sums                 : double[3] (sum|wl|, sum|dens|, gradSum)
 *   hpwl                 : int64[1]
 */
typedef struct rp_eval_out {
  int64_t* bin_inst;
  int64_t* bin_filler;
  float* density;
  float* phi;
  float* ex;
  float* ey;
  int64_t* overflow_area;
  float* net_sums;
  int32_t* net_box;
  float* wl_grad;
  float* dens_grad;
  float* sum_grad;
  double* sums;
  int64_t* hpwl;
} rp_eval_out;

int rp_eval_gradient(rp_ctx* ctx, const float* cx, const float* cy, float density_penalty,
                     float wl_coef_x, float wl_coef_y, rp_eval_out* out);

/* Poisson solve alone: density[B] (y-major) -> phi, ex, ey (y-major). */
int rp_eval_poisson(rp_ctx* ctx, const float* density, float* phi, float* ex, float* ey);

/* ------------------------------------------------------------------------
 * Poisson solve without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It solves
 * FFT::doFFT (src/fft.cpp:112-177) like rp_eval_poisson, on the caller's
 * device buffers (y-major, j * nx + i): density -> phi (may be NULL), ex,
 * ey, for an nx x ny grid of bins of bin_size_x x bin_size_y DBU, queued on
 * `stream` (a hipStream_t; NULL = the null stream) and not waited for.
 * Its work memory is a fixed scratch array in device memory, one per GPU,
 * held from the first call on that GPU until the process ends
 * (RP_NOALLOC_SCRATCH_MIB); it calls no hipMalloc / hipFree.  Calls that
 * overlap on one GPU share that array and race: order them on one stream.
 * Returns 0, RP_ERR_ARG (a NULL buffer, a non-positive bin size, a side
 * that is not a power of two in [4, 2048], or more than
 * RP_NOALLOC_MAX_BINS bins), RP_ERR_NO_DEVICE or RP_ERR_HIP.
 * ---------------------------------------------------------------------- */
#define RP_NOALLOC_MAX_BINS (1 << 20)  /* 1024 x 1024 */
#define RP_NOALLOC_SCRATCH_MIB 21      /* scratch held per GPU (rounded up) */
int rp_poisson_solve_noalloc(const float* density, int32_t nx, int32_t ny, int32_t bin_size_x,
                             int32_t bin_size_y, float* phi, float* ex, float* ey, void* stream);

/* ------------------------------------------------------------------------
 * Placement service (no reference counterpart: the reference places one
 * design per process).  Many jobs share batched evaluation steps on one GPU:
 * jobs are admitted in arrival order against an HBM budget, set up on host
 * worker threads, run as continuous batches (each step evaluates every
 * running job once through per-kernel block tables, captured in a HIP graph)
 * and retired as they finish.  Each job's result is bit-identical to running
 * it alone with rp_nesterov_place.  Used for the batched and served forms.
 * ---------------------------------------------------------------------- */
typedef struct rp_server rp_server;

typedef struct rp_server_opts {
  int64_t hbm_budget_bytes;  /* admission budget for job device memory */
  int32_t max_active;        /* jobs set up or running at once */
  int32_t setup_threads;     /* host workers for the per-job setup */
  int32_t steps_per_poll;    /* batched steps per graph replay */
  int32_t use_graph;         /* 1: replay a captured HIP graph */
  int32_t autotune;          /* 1: time the tile shapes of the batched scatter and
                                wirelength kernels per job size bucket on the
                                bucket's first jobs and keep the fastest */
  int32_t batched_ip;        /* 2 (default): the initial place of jobs of more
                                than 8192 movable instances (the grid-wide
                                solver) runs in rounds of shared launches (two
                                driver threads), smaller jobs' on their set-up
                                worker; 1: every job's in the rounds; 0: every
                                job's on its set-up worker.  Each job's result
                                is the same in every mode */
} rp_server_opts;

/* Tile shapes the service uses for one job size bucket (rp_server_tuning).
 * Buckets by movable + filler cell count: < 2k, < 8k, < 32k, < 128k, larger.
 * Every shape gives bit-identical results; only the time differs. */
#define RP_NUM_SIZE_BUCKETS 5
typedef struct rp_tile_choice {
  int32_t bucket;
  int32_t tuned;             /* 1: measured on this bucket's jobs, 0: defaults */
  int32_t jobs_sampled;
  int32_t scatter_cells;     /* cells per fused step+scatter workgroup (256, 512, 1024) */
  int32_t wa_pins;           /* pins per wirelength workgroup (256, 512, 1024) */
  float scatter_us[3];       /* measured mean launch time of each candidate */
  float wa_us[3];
} rp_tile_choice;

typedef struct rp_job_result {
  rp_nesterov_stats stats;
  int32_t code;              /* 0 or REPL code, < 0 on setup errors */
  int32_t done;
  double arrival_s, admit_s, start_s, done_s;  /* seconds from rp_server_run start */
  int64_t device_bytes;
} rp_job_result;

typedef struct rp_server_stats {
  int64_t steps;             /* batched evaluation steps issued */
  int32_t rebuilds;          /* block-table / graph rebuilds */
  double wall_s;
} rp_server_stats;

void rp_server_opts_default(rp_server_opts* opts);
int rp_server_create(const rp_server_opts* opts, rp_server** out);
/* Deep-copies the netlist; the job becomes eligible arrival_s seconds after
 * rp_server_run starts (0 = immediately). */
int rp_server_submit(rp_server* s, const rp_netlist* netlist, const rp_params* params, double arrival_s,
                     int32_t* job);
/* Runs until every submitted job has finished. */
int rp_server_run(rp_server* s);
int rp_server_job(rp_server* s, int32_t job, rp_job_result* out);
int rp_server_job_positions(rp_server* s, int32_t job, int32_t* lx, int32_t* ly);
int rp_server_job_records(rp_server* s, int32_t job, rp_iter_record* out, int32_t capacity, int32_t* count);
int rp_server_get_stats(rp_server* s, rp_server_stats* out);
/* The tile shape of every size bucket (RP_NUM_SIZE_BUCKETS entries). */
int rp_server_tuning(rp_server* s, rp_tile_choice* out, int32_t capacity, int32_t* count);
/* Pins a bucket's tile shape (256, 512 or 1024 each) before rp_server_run,
 * e.g. to reuse an earlier tuning; the tuner then leaves the bucket alone. */
int rp_server_set_tiles(rp_server* s, int32_t bucket, int32_t scatter_cells, int32_t wa_pins);
const char* rp_server_last_error(rp_server* s);
void rp_server_destroy(rp_server* s);

#ifdef __cplusplus
}
#endif

#endif /* REPLACE_RP_CAPI_H */
