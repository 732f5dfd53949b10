// rp_dev_struct.h -- kernel argument block shared by rp_kernels.hip and rp_host.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rp_device.h"

// Device pointers and sizes of one placement problem (layout in rp_device.h).
struct RpDev {
  int nc, nn, np, nbx, nby, bsx, bsy;
  int mfma;  // Poisson on matrix cores: the bin kernel only produces rho
  int scatter_per;  // cells per thread of the spatial-order scatter (1 or 2)
  int fuse_scatter_wa;  // single instance: density scatter and wirelength in one launch
  int core_lx, core_ly, core_ux, core_uy;
  float td;
  float bar;          // minWireLengthForceBar
  float min_precond;  // minPreconditioner
  const int2* dhalf;
  const float* scale;
  const uint8_t* kind;
  const float* area;
  const float* npin;
  float2* slp;
  float2* sumg;
  float2* co;
  int2* ic;
  const int* sord;     // cells in spatial order (density scatter of the single-instance path)
  const int* cp_start;
  const int* cp_pin;   // cell's pins (reference ITerm order) -> device pin index
  float2* pgrad;       // per-pin WA gradient gMin - gMax, device pin order
  const int* nstart;
  int nwa;             // wirelength workgroups of the packing in use
  const int4* wa_rng;  // their net ranges [x, y) and pin ranges [z, w)
  int nwa_v[RP_WA_VARIANTS];              // the packings for every tile-shape variant
  const int4* wa_rng_v[RP_WA_VARIANTS];
  const int* pcell;
  const int2* ppos;
  int4* nbox;
  float4* nsum;
  unsigned long long* bin_inst;
  unsigned long long* bin_filler;
  const float* bin_npf;  // float(nonPlace area) per bin (the value the density formulas convert it to)
  int store_rho;         // bin kernel writes rho (stage-level outputs, matrix-core solve)
  float* rho;
  float2* field;
  float* phi;
  float* t;
  float* y1;
  float* y2;
  const float2* twx;   // e^{-2 pi i k / Nx}, k < Nx/2
  const float2* tw2x;  // e^{-i pi k / (2 Nx)}, k < Nx
  const float2* twy;
  const float2* tw2y;
  const float* wx;
  const float* wx2;
  const float* wy;
  const float* wy2;
  double* part_grad;
  double* part_bin;
  long long* part_net;
  RpState* st;
  int npart_grad, npart_bin;  // partial sums per evaluation (RpState's num_partials_grad / _bin)
  RpIterRec* rec;
  float2* dbg_wl;
  float2* dbg_dens;
  // verbose-3 trace (all NULL unless the context traces): per-eval records,
  // per-block partials of the gradient's squared norm and of sum phi * area,
  // the per-bin area the bin kernel leaves for the latter
  RpEvalRec* evrec;
  int evrec_cap;
  double* part_gn;
  double* part_phi;
  float* dbg_area;
};

// Dynamic LDS bytes of the FFT kernels for an n-point row / column: two
// ping-pong buffers (n float2 plus one float2 of padding per 16) and the
// stage twiddles (n float2), plus the staged rows of the column and inverse
// passes.
inline size_t rp_lds_binfin(int n) { return sizeof(float2) * (2 * (size_t)(n + n / 16) + (size_t)n); }
inline size_t rp_lds_col(int n) { return rp_lds_binfin(n) + sizeof(float) * 4 * (size_t)n; }
inline size_t rp_lds_rinv(int n) { return rp_lds_binfin(n) + sizeof(float) * 4 * (size_t)n; }

// One batched evaluation step over many instances (kb_* kernels): per kernel
// a block table of (instance, logical block) entries (instance -1 = idle).
struct RpBatchLaunch {
  const RpDev* devs;
  const int2* tab[RP_NUM_EVAL_KERNELS];
  int nblk[RP_NUM_EVAL_KERNELS];
  const int2* tab_sc[RP_SC_VARIANTS];  // scatter and wirelength: one table per tile shape
  int nblk_sc[RP_SC_VARIANTS];
  const int2* tab_wa[RP_WA_VARIANTS];
  int nblk_wa[RP_WA_VARIANTS];
  size_t lds_binfin, lds_col, lds_rinv;
};

// Handle of the matrix-core Poisson solve (rp_dct_mfma.hip): its GEMM plans,
// split-BF16 transform matrices and intermediates.
struct RpMfmaDev {
  void* plan;
};

hipError_t rp_launch_poisson_mfma(const RpDev& d, const RpMfmaDev& m, hipStream_t s, int want_phi);
// One batched launch of tile-shape variant `variant` of kernel RP_K_SCATTER or
// RP_K_WA over block table tab (nblk entries).
hipError_t rp_launch_batch_variant(int kernel, int variant, const RpDev* devs, const int2* tab, int nblk,
                                   hipStream_t s);

// Work arrays of the spatial cell order (single-instance scatter): a
// counting sort of the cells by the tile holding their integer centre.
#define RP_SORT_TILES 16384  // 128 x 128 tiles at most
struct RpSortBufs {
  int* cnt;        // RP_SORT_TILES tile counts (zero between refreshes)
  int* off;        // RP_SORT_TILES exclusive offsets
  int* rank;       // Nc rank of each cell within its tile
  int* sord;       // Nc cells in tile order (RpDev::sord)
  int tile_shift;  // log2 of the tile side in bins
};

// Tile side (log2, in bins) so that an nbx x nby grid has <= 128 tiles per side.
int rp_sort_tile_shift(int nbx, int nby);
// Finishes the counting sort started by the last counting step launch:
// offsets, then d.sord.
hipError_t rp_launch_sort_refresh(const RpDev& d, const RpSortBufs& b, hipStream_t s);

// Load the code objects of rp_kernels.hip / rp_ip.hip on the current device
// (before many host threads start launching their kernels).
hipError_t rp_kernels_load_code();
hipError_t rp_ip_load_code();
