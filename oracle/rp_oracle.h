/*
 * rp_oracle.h -- CPU restatement of RePlAce's global-placement hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (replace_amd/) never links,
 * loads or calls it.
 *
 * It restates, single-threaded and with the reference's float / double /
 * long double promotions, the reference (The-OpenROAD-Project/RePlAce) files
 *   src/placerBase.cpp      PlacerBase::init, dummies, hpwl
 *   src/initialPlace.cpp    B2B model (solver: own BiCGSTAB, see below)
 *   src/nesterovBase.cpp    GCell/GPin/GNet/Bin/BinGrid, fillers, scatter,
 *                           WA wirelength, gradients
 *   src/nesterovPlace.cpp   init, gradients, back-tracking loop, schedules
 *   src/fft.cpp             Poisson wrapper (transform: pluggable, below)
 * Each function cites the reference lines it follows.
 *
 * Poisson transform.  The default is the oracle's own FP64 FFT-based
 * DCT-II / DCT-III / DST-III (Makhoul 1980) with the reference's FP32
 * rounding at the boundaries (inputs and outputs are float).  A caller may
 * plug in another solver through ro_set_poisson (the tests plug in the
 * reference's own Ooura transform, compiled from its sources into
 * oracle/_ref/, when that build is present).
 *
 * Parity pins (what fixes each part; see tests/):
 *   - scatter / bins / nonPlace / fillers / density sizes: brute force and
 *     closed forms on small grids, the reference's .ok traces.
 *   - Poisson: FP64 direct sums (closed form) and the reference Ooura build.
 *   - WA / gradients / step / schedules: closed forms (2-pin nets, finite
 *     differences), the reference's .ok traces (end to end).
 *   - BiCGSTAB: parity unpinned at the iterate level (Eigen is not in this
 *     image); pinned by residual invariants and the .ok initial-place HPWL.
 */
#ifndef RP_ORACLE_H
#define RP_ORACLE_H

#include <stdint.h>
#include "../include/replace/rp_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ro_ctx ro_ctx;

/* Poisson solver hook: rho is y-major [ny][nx] (bin index j*nx+i); outputs
 * likewise.  Must reproduce FFT::doFFT (src/fft.cpp:112-177). */
typedef void (*ro_poisson_fn)(void* user, int nx, int ny, int bsx, int bsy,
                              const float* rho, float* phi, float* ex, float* ey);

ro_ctx* ro_create(const rp_netlist* nl, const rp_params* p, int* code);
void ro_destroy(ro_ctx* c);
void ro_set_poisson(ro_ctx* c, ro_poisson_fn fn, void* user);
int ro_set_verbose_trace(ro_ctx* c, int on);
/* Seconds spent per stage of the Nesterov loop since the context was made
 * (instrumentation for the CPU-baseline split; it changes no result):
 * step + clamp, centres + density scatter + bin density, Poisson solve,
 * WA sums, cell gradients + preconditioner, step length. */
enum { RO_T_STEP = 0, RO_T_SCATTER, RO_T_POISSON, RO_T_WA, RO_T_GRAD, RO_T_STEPLEN, RO_STAGES };
int ro_get_stage_times(ro_ctx* c, double* out);

int ro_initial_place(ro_ctx* c, float* err, int64_t* hpwl, int cap, int* n_iter);

int ro_nesterov_setup(ro_ctx* c, rp_nb_info* info);
int ro_get_gcells(ro_ctx* c, int32_t* dhalf_x, int32_t* dhalf_y, float* scale,
                  int32_t* dx, int32_t* dy, float* init_cx, float* init_cy, uint8_t* kind);
int ro_get_nonplace(ro_ctx* c, int64_t* nonplace);
int ro_eval_gradient(ro_ctx* c, const float* cx, const float* cy, float density_penalty,
                     float wl_coef_x, float wl_coef_y, rp_eval_out* out);
int ro_eval_poisson(ro_ctx* c, const float* density, float* phi, float* ex, float* ey);

/* NesterovPlace::init + doNesterovPlace.  max_iters < 0: params value. */
int ro_nesterov_place(ro_ctx* c, rp_nesterov_stats* st);
/* Init only / bounded number of outer iterations (for timing samples). */
int ro_nesterov_init(ro_ctx* c);
int ro_nesterov_iterate(ro_ctx* c, int n_iters, rp_nesterov_stats* st);
int ro_get_iter_records(ro_ctx* c, rp_iter_record* out, int32_t cap, int32_t* count);
/* Loop state between two outer iterations (rp_nesterov_state layout). */
int ro_get_state(ro_ctx* c, rp_nesterov_state* s);
int ro_set_state(ro_ctx* c, const rp_nesterov_state* s);
int ro_get_positions(ro_ctx* c, int32_t* lx, int32_t* ly);
int ro_set_positions(ro_ctx* c, const int32_t* lx, const int32_t* ly);
int64_t ro_inst_hpwl(ro_ctx* c);
int ro_num_insts(ro_ctx* c);

/* Standalone transforms for tests: FP64 direct sums (reference definitions,
 * src/fftsg2d.cpp:200-333) and the oracle's own FFT Poisson. */
void ro_poisson_fft64(void* user, int nx, int ny, int bsx, int bsy,
                      const float* rho, float* phi, float* ex, float* ey);
void ro_poisson_direct64(void* user, int nx, int ny, int bsx, int bsy,
                         const float* rho, double* phi, double* ex, double* ey);

/* MT19937 (std::mt19937) draws, for the filler pin test. */
void ro_mt19937(uint32_t seed, uint32_t* out, int n);
float ro_fast_exp(float a);

#ifdef __cplusplus
}
#endif
#endif
