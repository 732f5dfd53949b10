// rp_device.h -- data shared between the host driver and the gfx950 kernels.
//
// Device memory layout (all SoA, one allocation per array):
//   gcells   (Nc = place insts + fillers, reference order of
//             NesterovBase::gCells_, src/nesterovBase.cpp:916-961)
//     dhalf  int2   half density size (dDx/2, dDy/2)
//     scale  float  densityScale
//     kind   uint8  0 std cell, 1 movable macro, 2 filler
//     area   float  float(dx)*float(dy) (density preconditioner)
//     npin   float  #pins (wire-length preconditioner)
//     slp    float2 [3][Nc] ring: cur / next / prev SLP coordinates
//     sumg   float2 [3][Nc] ring of preconditioned sum gradients
//     co     float2 [2][Nc] ring: cur / next coordinates
//     ic     int2   integer density centre of the last evaluated iterate
//     cpin   CSR (cell -> pins, reference ITerm order): net id + offset
//   nets     (Nn, renumbered by degree so a wavefront sees equal work;
//             pins of a net keep their reference order)
//     nstart int    pin range
//     pcell  int    gcell of the pin or -1 (fixed instance / port)
//     ppos   int2   offset from the cell centre, or absolute position
//     nbox   int4   pin bbox of the last evaluation
//     nsum   float4 [2][Nn]  (minX, xminX, maxX, xmaxX) (minY, yminY, maxY, ymaxY)
//   bins     (B = Nx*Ny, y-major j*Nx+i as BinGrid::bins_)
//     inst, filler, nonplace  int64 (scatter targets, unit DBU^2)
//     rho    float, field float2 (Ex, Ey), phi float (tests only)
//     FFT work arrays (Ny x Nx, row-major): t (row DCT), y1 / y2 (column pass)
#pragma once

#include <stdint.h>

#define RP_BLOCK 256
#define RP_CTRL_BLOCK 1024  // controller workgroup (fixed-order reductions)
// spatial-order scatter: two cells per thread from this many cells on (half
// the workgroups flush each dense bin), one below (enough workgroups to fill
// the GPU).  RP_SCATTER_PER=1|2 in the environment overrides the choice for
// contexts set up afterwards (both shapes give identical integer sums).
#define RP_SS_WIDE_CELLS (1 << 17)
#define RP_MAX_FFT 2048
#ifndef RP_WA_CAP
#define RP_WA_CAP 512         // pins per wirelength workgroup (a larger net gets one of its own)
#endif
#define RP_WA_NETS RP_BLOCK   // nets per wirelength workgroup
// Tile-shape variants of the batched launches (picked per size bucket by the
// placement service's tuner): cells per fused step+scatter workgroup
// 256 / 512 / 1024, pins per wirelength workgroup 256 / 512 / 1024.
#define RP_SC_VARIANTS 3
#define RP_WA_VARIANTS 3
#define RP_WA_DEFAULT_VARIANT 1  // 512 pins (RP_WA_CAP), the single-instance launch

// Kernels of one evaluation slot, in launch order (per-kernel timing index).
enum {
  RP_K_STEP = 0,  // single instance: step + integer centres (batched: fused into RP_K_SCATTER)
  RP_K_SCATTER,   // density scatter (single instance: spatial cell order)
  RP_K_PIN_POS,   // reserved (0): the wirelength kernel computes the pin centres as it loads them
  RP_K_BINFIN_ROWDCT,
  RP_K_MFMA_POISSON,  // matrix-core Poisson solve (replaces the next two)
  RP_K_COL_SPECTRAL,
  RP_K_ROW_INVERSE,
  RP_K_WA,
  RP_K_GRAD,
  RP_K_CONTROL,
  RP_NUM_EVAL_KERNELS
};

// Ring slots of the SLP / gradient buffers.
struct RpRing {
  int cur, next, prev;  // slp / sumg slots
  int co_cur, co_next;  // coordinate slots
};

// Scalar state of NesterovPlace (src/nesterovPlace.h:76-127) plus the loop
// locals of doNesterovPlace (src/nesterovPlace.cpp:300-313), kept in device
// memory so back-tracking and the schedules run without host round trips.
struct RpState {
  // ---- constants (set once) ----
  int num_cells;
  int max_iter;
  int max_back_track;
  int num_partials_grad;
  int num_partials_bin;
  int num_partials_net;
  float target_overflow;
  float min_phi_coef;
  float reference_hpwl;
  float init_density_penalty;
  float init_wl_coef;
  float area_denom;      // float(stdInstsArea) + float(macroInstsArea)*td
  float bin_size_sum;    // float(binSizeX + binSizeY)
  float initial_prev_coef;

  // ---- evolving ----
  int phase;             // 0 running, 1 finished, 2 diverged/stopped, 3 iteration cap
  int code;              // REPL code on divergence (0, 2, 3, 4)
  int iter;              // outer iteration index i (0-based)
  int bt;                // back-track index within iteration
  int evals;
  int mode;              // 0 step from ring.cur into ring.next; 1 evaluate slot `dst` as is
  int dst;               // slot being evaluated
  RpRing ring;
  float step_length;
  float coeff;
  float cur_a;
  float prev_a;          // a_k of the iteration in progress (cur_a is a_(k+1))
  float density_penalty;
  float wl_coef_x, wl_coef_y;
  float base_wl_coef;
  float sum_overflow;
  float min_sum_overflow;
  float hpwl_with_min_sum_overflow;
  float max_phi_coef;
  int is_max_phi_coef_changed;
  int diverged;
  long long prev_hpwl;
  // ---- last evaluation (reduced by the controller) ----
  double wl_grad_sum, dens_grad_sum, grad_sum;
  double coord_dist2, grad_dist2;
  double overflow_area_d;
  long long overflow_area;
  long long hpwl;
  float last_new_step;
};

// Partial-sum layout of the gradient kernel (per block).
enum {
  RP_P_WL = 0,
  RP_P_DENS = 1,
  RP_P_GRAD = 2,
  RP_P_DCOORD = 3,
  RP_P_DGRAD = 4,
  RP_P_COUNT = 5
};

// One gradient evaluation's values for the verbose-3 trace (the layout of
// rp_eval_record in rp_capi.h): written by the controller when the context
// traces.
struct RpEvalRec {
  int iter;   // outer iteration (0-based); -1 for the two NesterovPlace::init evaluations
  int trial;  // back-tracking trial within the iteration (init: 0 = at x0, 1 = at the previous SLP)
  int diverged;
  float density_penalty, wl_coef;
  float wl_grad_sum, density_grad_sum, grad_sum;
  float coordinate_distance, gradient_distance, new_step_length;
  float gradient_norm, sum_phi, overflow;
  long long hpwl;
};

struct RpIterRec {
  float overflow;
  float step_length;
  float density_penalty;
  int num_back_track;
  long long hpwl;
};
