// rp_kernels.hip -- gfx950 kernels of one Nesterov gradient evaluation and
// its on-device controller.
//
// One evaluation (reference src/nesterovPlace.cpp:338-377) is the launch
// sequence
//   k_step              step + clamp + int centre (A14, A5)
//   k_scatter           density scatter in spatial cell order (A5, A6)
//   k_binfin_rowdct     bin density, overflow partials, DCT-II along x (A6 tail, A8)
//   k_col_spectral      DCT-II along y, spectral divide, DCT-III / DST-III along y (A8)
//   k_row_inverse       DST-III / DCT-III along x -> field (Ex, Ey) (A8/A9)
//   k_wa                pin centres, net bbox, WA exponent sums, HPWL partials (A5, A10, A16)
//   k_grad              WL gradient, field gather, preconditioner, partial sums (A11-A13, A15)
//   k_control_main      reductions + back-tracking / schedules / stop test (A15, A16, A18)
// Every kernel reads the controller state first and returns when the run
// has stopped, so the host can queue evaluation slots (or replay a graph of
// them) without knowing how many back-tracking trials an iteration takes.
//
// Arithmetic follows the reference's float / double promotions and must be
// built with -ffp-contract=off (no FMA contraction).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include "rp_device.h"

#include "rp_dev_struct.h"

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// fastExp, reference src/nesterovBase.cpp:1568-1582.  Its first step,
// (float)(1.0 + (double)a / 1024.0), is computed in float: a / 1024 is exact,
// and the double sum is exact whenever it can differ from 1.0f, so one float
// rounding gives the same value (checked for every finite float).  As one
// FMA: the product by 2^-10 is exact (or, below 2^-116, too small to move
// 1.0f), so fma(a, 2^-10, 1) rounds the same sum once, like the add.
__device__ __forceinline__ float fast_exp(float a) {
  a = __builtin_fmaf(a, 0.0009765625f, 1.0f);
  a = a * a; a = a * a; a = a * a; a = a * a; a = a * a;
  a = a * a; a = a * a; a = a * a; a = a * a; a = a * a;
  return a;
}

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Fixed-order block sum (deterministic); result valid in thread 0.
__device__ double block_sum_d(double v, double* sh) {
  v = wave_sum_d(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
  __syncthreads();
  return r;
}

__device__ long long block_sum_ll(long long v, long long* sh) {
  v = wave_sum_ll(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  long long r = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
  __syncthreads();
  return r;
}

// Bin geometry of BinGrid::initBins (src/nesterovBase.cpp:674-711): regular
// bins of binSize, the last column / row truncated at the core edge.
__device__ __forceinline__ void bin_x(const RpDev& d, int i, int& lx, int& ux) {
  lx = d.core_lx + i * d.bsx;
  ux = min(lx + d.bsx, d.core_ux);
}
__device__ __forceinline__ void bin_y(const RpDev& d, int j, int& ly, int& uy) {
  ly = d.core_ly + j * d.bsy;
  uy = min(ly + d.bsy, d.core_uy);
}

// getDensityMinMaxIdxX/Y (src/nesterovBase.cpp:822-841), clamped to the grid
__device__ __forceinline__ void idx_range(int lo_c, int hi_c, int core_l, int bs, int cnt, int& a,
                                          int& b) {
  a = (lo_c - core_l) / bs;
  int r = hi_c - core_l;
  int fm = r >= bs ? r % bs : r;
  b = (fm == 0) ? r / bs : r / bs + 1;
  a = max(a, 0);
  b = min(b, cnt);
}

// getDensityCoordiLayoutInsideX/Y (src/nesterovBase.cpp:1270-1297)
__device__ __forceinline__ float clamp_in(float c, int h, int lo, int hi) {
  float v = c;
  if (c - (float)h < (float)lo) v = (float)(lo + h);
  if (c + (float)h > (float)hi) v = (float)(hi - h);
  return v;
}

// XCD-aware block order.  MI355X deals workgroups round-robin to its 8 XCDs,
// each with a private L2.  Hardware block b runs on XCD b % 8; this maps it
// to logical block (b % 8) * per + min(b % 8, rem) + b / 8, so every XCD
// walks a contiguous range of logical blocks (neighbouring cells / nets in
// the locality order share its L2).  A bijection on [0, nblocks).
#define RP_NUM_XCD 8
// The run state's phase through a vector load: a scalar load of it would be
// waited for together with the kernel's later scalar loads (arguments), which
// puts its memory round trip in front of the kernel's first data loads.
__device__ __forceinline__ int run_phase(const RpState* st) {
  // a relaxed atomic load is never made scalar, and is not waited for at once
  // (the state was written by an earlier kernel: the launch boundary orders it)
  return __hip_atomic_load(&st->phase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  int per = nblocks / RP_NUM_XCD, rem = nblocks % RP_NUM_XCD;
  int x = b % RP_NUM_XCD, j = b / RP_NUM_XCD;
  return x * per + min(x, rem) + j;
}

// ---------------------------------------------------------------------------
// Stockham FFT in LDS, specialised per transform length N (natural order in,
// natural order out): radix-4 stages, preceded by one radix-2 stage when
// log2(N) is odd, so a 1024-point transform takes 5 barriers.  The work
// arrays are indexed through lpad (one float2 of padding per 16 entries),
// which spreads the power-of-two strides of the stage reads and writes over
// the LDS banks.  The twiddles of every radix-4 stage with ns > 1 sit in a
// table of their own (w1, w2, w3 for k < ns, stage after stage), so a
// stage reads them contiguously.  DIR = -1 forward (e^{-i}), +1 inverse
// (e^{+i}), unnormalised.  The transform's products may contract to FMAs:
// its parity bar is a tolerance against the reference's Ooura transform.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }

// float2 entries of one padded work array of n points
__host__ __device__ constexpr int fft_buf(int n) { return n + (n >> 4); }

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

// z * (dir * i)
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 z) {
  return DIR > 0 ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

// Stage twiddle tables of an N-point transform from the global half circle
// tw[m] = e^{-2 pi i m / N}, m < N/2 (e^{-2 pi i m / N} = -tw[m - N/2] above):
// for every radix-4 stage with ns > 1, w^r for r = 1..3 and k < ns (entry
// (r - 1) ns + k of the stage's table, the stages' tables one after the
// other).  In two steps: the global loads into registers (NT threads, entry
// g = tid + r NT), then the LDS stores, so the loads go out beside the
// kernel's data loads with no wait between them.
constexpr int stage_tw_entries(int n) {
  int e = 0;
  for (int ns = (ilog2c(n) & 1) ? 2 : 4; ns < n; ns *= 4) e += 3 * ns;
  return e;
}
template <int N>
__device__ __forceinline__ int stage_tw_m(int g) {  // the circle index m of entry g
  constexpr int LOG = ilog2c(N);
  int off = 0, m = 0;
#pragma unroll
  for (int ns = (LOG & 1) ? 2 : 4; ns < N; ns *= 4) {
    if (g >= off && g < off + 3 * ns) {
      const int e = g - off, r = e / ns + 1, k = e - (r - 1) * ns;
      m = r * k * (N / (4 * ns));
    }
    off += 3 * ns;
  }
  return m;
}
template <int N, int NT>
struct StageTw {
  static constexpr int E = stage_tw_entries(N);
  static constexpr int R = (E + NT - 1) / NT;
  float2 v[R > 0 ? R : 1];
  unsigned neg = 0;  // bit r: entry r is the negated circle value
  // unconditional loads (an entry past the tables repeats the last one and
  // is not stored); the sign is applied at the store
  __device__ __forceinline__ void load(const float2* __restrict__ tw, int tid) {
    if constexpr (E > 0) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int m = stage_tw_m<N>(min(tid + r * NT, E - 1));
        v[r] = tw[m < N / 2 ? m : m - N / 2];
        neg |= (m >= N / 2 ? 1u : 0u) << r;
      }
    }
  }
  __device__ __forceinline__ void store(float2* TW, int tid) const {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int g = tid + r * NT;
      if (g < E) TW[g] = (neg >> r & 1) ? make_float2(-v[r].x, -v[r].y) : v[r];
    }
  }
};

// One radix-4 stage of stride ns (twiddles at TW + off), in -> out.
template <int N, int NS, int DIR>
__device__ __forceinline__ void fft_stage4(const float2* in, float2* out, const float2* TW, int off, int tid,
                                           int nt) {
#pragma clang fp contract(fast)
  constexpr int Q = N / 4;
  for (int j = tid; j < Q; j += nt) {
    const int k = j & (NS - 1);
    float2 a0 = in[lpad(j)], a1 = in[lpad(j + Q)], a2 = in[lpad(j + 2 * Q)], a3 = in[lpad(j + 3 * Q)];
    if (NS > 1) {
      float2 w1 = TW[off + k], w2 = TW[off + NS + k], w3 = TW[off + 2 * NS + k];
      if (DIR > 0) {
        w1.y = -w1.y;
        w2.y = -w2.y;
        w3.y = -w3.y;
      }
      auto cm = [](float2 x, float2 w) { return make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x); };
      a1 = cm(a1, w1);
      a2 = cm(a2, w2);
      a3 = cm(a3, w3);
    }
    const float2 s02 = make_float2(a0.x + a2.x, a0.y + a2.y);
    const float2 d02 = make_float2(a0.x - a2.x, a0.y - a2.y);
    const float2 s13 = make_float2(a1.x + a3.x, a1.y + a3.y);
    const float2 d13 = mul_dir_i<DIR>(make_float2(a1.x - a3.x, a1.y - a3.y));
    const int dst = (j - k) * 4 + k;
    out[lpad(dst)] = make_float2(s02.x + s13.x, s02.y + s13.y);
    out[lpad(dst + NS)] = make_float2(d02.x + d13.x, d02.y + d13.y);
    out[lpad(dst + 2 * NS)] = make_float2(s02.x - s13.x, s02.y - s13.y);
    out[lpad(dst + 3 * NS)] = make_float2(d02.x - d13.x, d02.y - d13.y);
  }
}

template <int N, int NS, int DIR>
__device__ __forceinline__ float2* fft_stages(float2* in, float2* out, const float2* TW, int off, int tid, int nt) {
  if constexpr (NS >= N) {
    return in;
  } else {
    fft_stage4<N, NS, DIR>(in, out, TW, off, tid, nt);
    __syncthreads();
    return fft_stages<N, NS * 4, DIR>(out, in, TW, NS > 1 ? off + 3 * NS : off, tid, nt);
  }
}

// N-point transform of a (b is the other work array) by threads tid of nt;
// returns the array holding the result.  Every thread of the workgroup
// reaches its barriers (groups of threads may run transforms of the same
// length side by side on their own arrays).
template <int N, int DIR>
__device__ __forceinline__ float2* fft_n(float2* a, float2* b, const float2* TW, int tid, int nt) {
  constexpr int LOG = ilog2c(N);
  if constexpr (LOG & 1) {  // radix-2 first stage (no twiddles)
    constexpr int H = N / 2;
    for (int j = tid; j < H; j += nt) {
      const float2 u = a[lpad(j)], v = a[lpad(j + H)];
      b[lpad(2 * j)] = make_float2(u.x + v.x, u.y + v.y);
      b[lpad(2 * j + 1)] = make_float2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
    return fft_stages<N, 2, DIR>(b, a, TW, 0, tid, nt);
  } else {
    return fft_stages<N, 1, DIR>(a, b, TW, 0, tid, nt);
  }
}

template <int N, int DIR>
__device__ __forceinline__ float2* fft_n(float2* a, float2* b, const float2* TW) {
  return fft_n<N, DIR>(a, b, TW, threadIdx.x, blockDim.x);
}

// Makhoul input reorder index: x[i] goes to v[pos(i)]
__device__ __forceinline__ int mk_pos(int i, int n) { return (i & 1) ? (n - 1 - (i >> 1)) : (i >> 1); }

// After a forward FFT of z = v_a + i v_b (each Makhoul-reordered), the two
// DCT-II spectra X[k] = sum_n x[n] cos(pi (n+1/2) k / N), with the twiddle
// tw2[k] = (cos t, -sin t), t = pi k / 2N, already loaded (w).
template <int N>
__device__ __forceinline__ void dct2_post(const float2* Z, int k, float2 w, float& xa, float& xb) {
  float2 zk = Z[lpad(k)];
  float2 zm = Z[lpad((N - k) & (N - 1))];
  zm.y = -zm.y;  // conj
  const float2 va = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y));
  const float dr = zk.x - zm.x, di = zk.y - zm.y;
  const float2 vb = make_float2(0.5f * di, -0.5f * dr);
  xa = va.x * w.x - va.y * w.y;
  xb = vb.x * w.x - vb.y * w.y;
}

// Pre-twiddle of the inverse Makhoul (DCT-III, Ooura "IDCT" without scale):
// W[k] = 1/2 e^{i pi k/2N} (a[k] - i a[N-k]) (k >= 1), W[0] = a[0];
// returns W_a + i W_b for two real coefficient rows, with the twiddle
// tw2[k] already loaded (w; unused at k = 0).
__device__ __forceinline__ float2 dct3_pre(float a_k, float a_nk, float b_k, float b_nk, int k, float2 w) {
  float2 wa, wb;
  if (k == 0) {
    wa = make_float2(a_k, 0.0f);
    wb = make_float2(b_k, 0.0f);
  } else {
    const float c = w.x, s = -w.y;  // e^{+i t} = (c, s)
    wa = make_float2(0.5f * (a_k * c + a_nk * s), 0.5f * (a_k * s - a_nk * c));
    wb = make_float2(0.5f * (b_k * c + b_nk * s), 0.5f * (b_k * s - b_nk * c));
  }
  return make_float2(wa.x - wb.y, wa.y + wb.x);
}

// Output m of an inverse Makhoul transform held in Z: y[2q] = w[q],
// y[2q+1] = w[N-1-q]
template <int N>
__device__ __forceinline__ float2 dct3_out(const float2* Z, int m) {
  return Z[lpad((m & 1) ? (N - 1 - (m >> 1)) : (m >> 1))];
}

// ---------------------------------------------------------------------------
// 1. step + clamp + integer centre + density scatter
// ---------------------------------------------------------------------------
// LDS aggregation: each workgroup scatters a run of cells into a hashed
// table of bins in LDS (linear probing, RP_AGG_PROBES slots) and flushes one
// int64 atomic per touched bin; a bin that finds no free slot goes straight
// to global memory.  Integer sums are order independent, so the result is
// exact and deterministic whatever order the cells come in.
#ifndef RP_AGG_PROBES
#define RP_AGG_PROBES 4
#endif

template <int BITS>
struct AggTable {
  static constexpr int SLOTS = 1 << BITS;
  int key[SLOTS];
  unsigned long long inst[SLOTS];
  unsigned long long fill[SLOTS];
};

template <int BITS>
__device__ __forceinline__ void agg_clear(AggTable<BITS>& t) {
  for (int k = threadIdx.x; k < AggTable<BITS>::SLOTS; k += blockDim.x) {
    t.key[k] = -1;
    t.inst[k] = 0;
    t.fill[k] = 0;
  }
}

template <int BITS>
__device__ __forceinline__ void agg_add(const RpDev& d, AggTable<BITS>& t, int b, unsigned long long a, bool filler) {
  int slot = (int)(((unsigned)b * 2654435761u) >> (32 - BITS));
#pragma unroll
  for (int q = 0; q < RP_AGG_PROBES; q++) {
    int sl = (slot + q) & (AggTable<BITS>::SLOTS - 1);
    int old = atomicCAS(&t.key[sl], -1, b);
    if (old == -1 || old == b) {
      atomicAdd(filler ? &t.fill[sl] : &t.inst[sl], a);
      return;
    }
  }
  atomicAdd(filler ? &d.bin_filler[b] : &d.bin_inst[b], a);
}

template <int BITS>
__device__ __forceinline__ void agg_flush(const RpDev& d, const AggTable<BITS>& t) {
  for (int k = threadIdx.x; k < AggTable<BITS>::SLOTS; k += blockDim.x) {
    int b = t.key[k];
    if (b >= 0) {
      if (t.inst[k]) atomicAdd(&d.bin_inst[b], t.inst[k]);
      if (t.fill[k]) atomicAdd(&d.bin_filler[b], t.fill[k]);
    }
  }
}

// Window table for the spatially ordered scatter: bin (i, j) maps to slot
// (j mod RP_WIN_Y) * RP_WIN_X + (i mod RP_WIN_X), a tiling of the plane, so
// the compact region a workgroup's cells cover lands on distinct slots and
// the flush, walking the slots in order, issues atomics to consecutive bins
// of a row (coalesced).  A slot held by another bin is probed past in
// steps of 7; a bin that finds no slot goes straight to global memory.
#ifndef RP_WIN_X
#define RP_WIN_X 32
#endif
#ifndef RP_WIN_Y
#define RP_WIN_Y 32
#endif
struct WinTable {
  static constexpr int SLOTS = RP_WIN_X * RP_WIN_Y;
  int key[SLOTS];
  unsigned long long inst[SLOTS];
  unsigned long long fill[SLOTS];
};

__device__ __forceinline__ void win_clear(WinTable& t) {
  for (int k = threadIdx.x; k < WinTable::SLOTS; k += blockDim.x) {
    t.key[k] = -1;
    t.inst[k] = 0;
    t.fill[k] = 0;
  }
}

__device__ __forceinline__ void win_add(const RpDev& d, WinTable& t, int i, int j, unsigned long long a, bool filler) {
  const int b = j * d.nbx + i;
  const int s0 = (j & (RP_WIN_Y - 1)) * RP_WIN_X + (i & (RP_WIN_X - 1));
#pragma unroll
  for (int q = 0; q < RP_AGG_PROBES; q++) {
    int sl = (s0 + 7 * q) & (WinTable::SLOTS - 1);
    int k = t.key[sl];
    if (k == -1) k = atomicCAS(&t.key[sl], -1, b);  // claim (returns the holder if another lane won)
    if (k == -1 || k == b) {
      atomicAdd(filler ? &t.fill[sl] : &t.inst[sl], a);
      return;
    }
  }
  atomicAdd(filler ? &d.bin_filler[b] : &d.bin_inst[b], a);
}

__device__ __forceinline__ void win_flush(const RpDev& d, const WinTable& t) {
  for (int k = threadIdx.x; k < WinTable::SLOTS; k += blockDim.x) {
    int b = t.key[k];
    if (b >= 0) {
      if (t.inst[k]) atomicAdd(&d.bin_inst[b], t.inst[k]);
      if (t.fill[k]) atomicAdd(&d.bin_filler[b], t.fill[k]);
    }
  }
}

// Nesterov step of one cell (src/nesterovPlace.cpp:339-361, no FMA: built
// with -ffp-contract=off) with the layout clamp of both results
// (getDensityCoordiLayoutInsideX/Y).  Writes the next coordinate and the
// next SLP; returns the SLP, the iterate that is evaluated.
__device__ __forceinline__ float2 step_cell(const RpDev& d, const RpState* st, int c, int2 h) {
  int cur = st->ring.cur;
  float2 x = d.slp[(size_t)cur * d.nc + c];
  float2 g = d.sumg[(size_t)cur * d.nc + c];
  float2 o = d.co[(size_t)st->ring.co_cur * d.nc + c];
  float sl = st->step_length, cf = st->coeff;
  float nx = x.x + sl * g.x;
  float ny = x.y + sl * g.y;
  float sx = nx + cf * (nx - o.x);
  float sy = ny + cf * (ny - o.y);
  float2 nco = make_float2(clamp_in(nx, h.x, d.core_lx, d.core_ux), clamp_in(ny, h.y, d.core_ly, d.core_uy));
  float2 p = make_float2(clamp_in(sx, h.x, d.core_lx, d.core_ux), clamp_in(sy, h.y, d.core_ly, d.core_uy));
  d.co[(size_t)st->ring.co_next * d.nc + c] = nco;
  d.slp[(size_t)st->dst * d.nc + c] = p;
  return p;
}

// The iterate one evaluation scatters: the stepped SLP (mode 0) or slot dst
// as it is (mode 1).  setDensityCenterLocation(int, int) truncates the float
// centre toward zero (src/nesterovBase.cpp:1203); the integer centre is kept
// for the pin, wirelength and field kernels.
__device__ __forceinline__ int2 eval_centre(const RpDev& d, const RpState* st, int c, int2 h) {
  float2 p = st->mode == 0 ? step_cell(d, st, c, h) : d.slp[(size_t)st->dst * d.nc + c];
  int2 icc = make_int2((int)p.x, (int)p.y);
  d.ic[c] = icc;
  return icc;
}

// getOverlapDensityArea * densityScale (* td for macros), truncated to int64,
// for every bin of the cell's density box (updateBinsGCellDensityArea,
// src/nesterovBase.cpp:745-799).
template <int BITS>
__device__ __forceinline__ void box_scatter(const RpDev& d, AggTable<BITS>& t, int2 icc, int2 h, float sc,
                                            uint8_t kd) {
  int dlx = icc.x - h.x, dux = icc.x + h.x, dly = icc.y - h.y, duy = icc.y + h.y;
  int x0, x1, y0, y1;
  idx_range(dlx, dux, d.core_lx, d.bsx, d.nbx, x0, x1);
  idx_range(dly, duy, d.core_ly, d.bsy, d.nby, y0, y1);
  for (int i = x0; i < x1; i++) {
    int blx, bux;
    bin_x(d, i, blx, bux);
    int w = min(bux, dux) - max(blx, dlx);
    if (w <= 0) continue;
    for (int j = y0; j < y1; j++) {
      int bly, buy;
      bin_y(d, j, bly, buy);
      int hh = min(buy, duy) - max(bly, dly);
      if (hh <= 0) continue;
      float v = (float)w * (float)hh;
      v = v * sc;
      if (kd == 1) v = v * d.td;
      unsigned long long a = (unsigned long long)(long long)v;
      if (a == 0) continue;
      agg_add(d, t, j * d.nbx + i, a, kd == 2);
    }
  }
}

__device__ __forceinline__ void box_scatter_win(const RpDev& d, WinTable& t, int2 icc, int2 h, float sc,
                                                uint8_t kd) {
  int dlx = icc.x - h.x, dux = icc.x + h.x, dly = icc.y - h.y, duy = icc.y + h.y;
  int x0, x1, y0, y1;
  idx_range(dlx, dux, d.core_lx, d.bsx, d.nbx, x0, x1);
  idx_range(dly, duy, d.core_ly, d.bsy, d.nby, y0, y1);
  for (int i = x0; i < x1; i++) {
    int blx, bux;
    bin_x(d, i, blx, bux);
    int w = min(bux, dux) - max(blx, dlx);
    if (w <= 0) continue;
    for (int j = y0; j < y1; j++) {
      int bly, buy;
      bin_y(d, j, bly, buy);
      int hh = min(buy, duy) - max(bly, dly);
      if (hh <= 0) continue;
      float v = (float)w * (float)hh;
      v = v * sc;
      if (kd == 1) v = v * d.td;
      unsigned long long a = (unsigned long long)(long long)v;
      if (a == 0) continue;
      win_add(d, t, i, j, a, kd == 2);
    }
  }
}

// Fused step + scatter over PER * RP_BLOCK consecutive cells into a
// 2^BITS-slot table (batched instances: small grids, where the locality
// order keeps a workgroup's bins together).  The batched step launches one
// of the RP_SC_VARIANTS tile shapes per size bucket (chosen by the service's
// tuner); the integer sums do not depend on the shape.
template <int PER, int BITS>
__device__ void scatter_body(const RpDev& d, int blk) {
  __shared__ AggTable<BITS> t;
  const RpState* st = d.st;
  if (st->phase != 0) return;
  agg_clear(t);
  __syncthreads();
  for (int r = 0; r < PER; r++) {
    int c = blk * PER * RP_BLOCK + r * RP_BLOCK + threadIdx.x;
    if (c >= d.nc) break;
    int2 h = d.dhalf[c];
    int2 icc = eval_centre(d, st, c, h);
    box_scatter(d, t, icc, h, d.scale[c], d.kind[c]);
  }
  __syncthreads();
  agg_flush(d, t);
}

// Tile key of an integer centre: Morton code of the (2^tile_shift)^2-bin
// tile holding it (at most 128 tiles per side, so keys < RP_SORT_TILES).
__device__ __forceinline__ unsigned spread7(unsigned v) {
  v &= 0x7fu;
  v = (v | (v << 4)) & 0x0f0fu;
  v = (v | (v << 2)) & 0x3333u;
  v = (v | (v << 1)) & 0x5555u;
  return v;
}
__device__ __forceinline__ int tile_key(const RpDev& d, int2 p, int tile_shift) {
  int bx = min(max((p.x - d.core_lx) / d.bsx, 0), d.nbx - 1) >> tile_shift;
  int by = min(max((p.y - d.core_ly) / d.bsy, 0), d.nby - 1) >> tile_shift;
  return (int)(spread7((unsigned)bx) | (spread7((unsigned)by) << 1));
}

// Step only, one thread per cell in storage order (coalesced): next
// coordinate, next SLP and the integer centre.
__device__ void step_body(const RpDev& d, int blk) {
  const RpState* st = d.st;
  if (st->phase != 0) return;
  int c = blk * RP_BLOCK + threadIdx.x;
  if (c >= d.nc) return;
  (void)eval_centre(d, st, c, d.dhalf[c]);
}

// Scatter only, over PER * RP_BLOCK cells of the spatial order d.sord
// (cells sorted by the Morton code of their bin tile), so a workgroup's
// cells share bins and nearly every update lands in the LDS table.  Any
// permutation gives the same integer sums; the order only decides how many
// updates go to global memory.
//
// With `count` set the kernel also starts the next refresh of the order (a
// counting sort by tile): each cell is counted into its tile, first in a
// small LDS table (the workgroup's cells cover few tiles), then with one
// global atomic per tile and workgroup, and its rank within the tile is
// kept in sb.rank.  rp_launch_sort_refresh turns counts and ranks into the
// new order.
#define RP_TILE_SLOTS 256
struct SortScatterLds {
  WinTable t;
  int key[RP_TILE_SLOTS], cnt[RP_TILE_SLOTS];  // cnt becomes the tile's base rank after the flush
};
template <int PER>
__device__ void sorted_scatter_body(const RpDev& d, int blk, const RpSortBufs& sb, int count, SortScatterLds& L) {
  WinTable& t = L.t;
  int* t_key = L.key;
  int* t_cnt = L.cnt;
  const RpState* st = d.st;
  if (st->phase != 0) return;
  int cell[PER], slot[PER], lrank[PER];
  // all loads before the first atomic (which could alias them), and before
  // the table clear and its barrier, which they overlap
  int2 v_ic[PER], v_h[PER];
  float v_sc[PER];
  uint8_t v_kd[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    int k = blk * (PER * RP_BLOCK) + r * RP_BLOCK + threadIdx.x;
    cell[r] = k < d.nc ? d.sord[k] : -1;
  }
#pragma unroll
  for (int r = 0; r < PER; r++) {
    int c = cell[r];
    if (c < 0) continue;
    v_ic[r] = d.ic[c];
    v_h[r] = d.dhalf[c];
    v_sc[r] = d.scale[c];
    v_kd[r] = d.kind[c];
  }
  win_clear(t);
  for (int k = threadIdx.x; k < RP_TILE_SLOTS; k += blockDim.x) {
    t_key[k] = -1;
    t_cnt[k] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; r++) {
    int c = cell[r];
    if (c < 0) continue;
    int2 icc = v_ic[r];
    box_scatter_win(d, t, icc, v_h[r], v_sc[r], v_kd[r]);
    if (count) {
      cell[r] = c;
      slot[r] = -1;
      int key = tile_key(d, icc, sb.tile_shift);
      int h = (int)(((unsigned)key * 2654435761u) >> 24);  // 8-bit hash
      for (int q = 0; q < 8; q++) {
        int sl = (h + q) & (RP_TILE_SLOTS - 1);
        int old = atomicCAS(&t_key[sl], -1, key);
        if (old == -1 || old == key) {
          slot[r] = sl;
          lrank[r] = atomicAdd(&t_cnt[sl], 1);
          break;
        }
      }
      if (slot[r] < 0) sb.rank[c] = atomicAdd(&sb.cnt[key], 1);  // table full: count directly
    }
  }
  __syncthreads();
  win_flush(d, t);
  if (count) {
    for (int k = threadIdx.x; k < RP_TILE_SLOTS; k += blockDim.x)
      if (t_key[k] >= 0) t_cnt[k] = atomicAdd(&sb.cnt[t_key[k]], t_cnt[k]);  // count -> base rank
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; r++)
      if (cell[r] >= 0 && slot[r] >= 0) sb.rank[cell[r]] = t_cnt[slot[r]] + lrank[r];
  }
}

// sorted_scatter_body with its own LDS
template <int PER>
__device__ void sorted_scatter_body(const RpDev& d, int blk, const RpSortBufs& sb, int count) {
  __shared__ SortScatterLds L;
  sorted_scatter_body<PER>(d, blk, sb, count, L);
}

// ---------------------------------------------------------------------------
// 1b. pin centres (GPin::updateDensityLocation, src/nesterovBase.cpp:411-415):
// integer density centre + offset; fixed and port pins (no gcell) keep their
// absolute position.  The wirelength kernel computes them as it loads its
// pins, so they are never stored.
__device__ __forceinline__ int2 pin_centre(const RpDev& d, int q) {
  const int c = d.pcell[q];
  int2 o = d.ppos[q];
  if (c >= 0) {
    const int2 icc = d.ic[c];
    o = make_int2(icc.x + o.x, icc.y + o.y);
  }
  return o;
}

// ---------------------------------------------------------------------------
// 2. bin density + overflow + DCT-II along x (two rows per workgroup)
// ---------------------------------------------------------------------------
// LDS of the FFT kernels (dynamic): the two padded work arrays and the stage
// twiddles (N float2 each at most), then the kernel's staged rows.
template <int N>
__device__ void binfin_body(const RpDev& d, int blk, int from_rho) {
  extern __shared__ float2 smem[];
  __shared__ double sh_red[RP_BLOCK / 64];
  // the run state is checked once the loads below are in flight (a stopped
  // run's kernels return before any store)
  float2* A = smem;
  float2* B = smem + fft_buf(N);
  float2* TW = smem + 2 * fft_buf(N);
  const int ja = 2 * blk, jb = ja + 1;
  double ov_sum = 0.0;
  constexpr int MAXC = (N + RP_BLOCK - 1) / RP_BLOCK;
  // the DCT-II post-twiddles of the thread's outputs k = threadIdx.x + r
  // blockDim.x, loaded now rather than after the FFT
  float2 tw[MAXC];
  StageTw<N, RP_BLOCK> stw;
  if (!d.mfma) {
#pragma unroll
    for (int r = 0; r < MAXC; r++) {
      const int k = threadIdx.x + r * blockDim.x;
      if (k < N) tw[r] = d.tw2x[k];
    }
    stw.load(d.twx, threadIdx.x);
  }
  if (from_rho) {  // Poisson-only entry: density given
    if (run_phase(d.st) != 0) return;
    if (!d.mfma) stw.store(TW, threadIdx.x);
    for (int i = threadIdx.x; i < N; i += blockDim.x)
      A[lpad(mk_pos(i, N))] = make_float2(d.rho[(size_t)ja * N + i], d.rho[(size_t)jb * N + i]);
  } else {
    // every load of the thread's bins is issued before any store (the zeroed
    // accumulators could otherwise alias later loads and serialise them)
    long long v_in[2][MAXC], v_fi[2][MAXC];
    float v_np[2][MAXC];
#pragma unroll
    for (int t = 0; t < MAXC; t++) {
      const int i = threadIdx.x + t * blockDim.x;
      if (i < N)
#pragma unroll
        for (int r = 0; r < 2; r++) {
          const size_t b = (size_t)(ja + r) * N + i;
          v_in[r][t] = (long long)d.bin_inst[b];
          v_fi[r][t] = (long long)d.bin_filler[b];
          v_np[r][t] = d.bin_npf[b];
        }
    }
    if (run_phase(d.st) != 0) return;
    if (!d.mfma) stw.store(TW, threadIdx.x);
#pragma unroll
    for (int t = 0; t < MAXC; t++) {
      const int i = threadIdx.x + t * blockDim.x;
      if (i >= N) continue;
      int blx, bux;
      bin_x(d, i, blx, bux);
      float r2[2];
#pragma unroll
      for (int r = 0; r < 2; r++) {
        int bly, buy;
        bin_y(d, ja + r, bly, buy);
        const size_t b = (size_t)(ja + r) * N + i;
        const long long bin_area = (long long)(bux - blx) * (long long)(buy - bly);
        const float at = (float)bin_area * d.td;
        // src/nesterovBase.cpp:806-816
        const float rho = ((float)v_in[r][t] + (float)v_fi[r][t] + v_np[r][t]) / at;
        float ov = (float)v_in[r][t] + v_np[r][t] - at;
        ov = (0.0f < ov) ? ov : 0.0f;
        ov_sum += (double)ov;
        if (d.store_rho) d.rho[b] = rho;
        if (d.dbg_area) d.dbg_area[b] = (float)(v_in[r][t] + v_fi[r][t]) + v_np[r][t];
        d.bin_inst[b] = 0ull;
        d.bin_filler[b] = 0ull;
        r2[r] = rho;
      }
      A[lpad(mk_pos(i, N))] = make_float2(r2[0], r2[1]);
    }
  }
  const double tot = block_sum_d(ov_sum, sh_red);
  if (threadIdx.x == 0) d.part_bin[blk] = tot;
  if (d.mfma) return;  // the matrix-core solve takes rho from d.rho
  // block_sum_d synchronised the LDS writes above
  const float2* Z = fft_n<N, -1>(A, B, TW);
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int k = threadIdx.x + r * blockDim.x;
    if (k >= N) continue;
    float xa, xb;
    dct2_post<N>(Z, k, tw[r], xa, xb);
    d.t[(size_t)ja * N + k] = xa;
    d.t[(size_t)jb * N + k] = xb;
  }
}

// ---------------------------------------------------------------------------
// 4. column pass (u fixed, v along the column), read from and written to
// the row-major grids: DCT-II along y, FFT::doFFT's scaling and spectral
// divide (src/fft.cpp:117-166), then DCT-III (phi) and DST-III (Ey) along y.
// ---------------------------------------------------------------------------
template <int N>
__device__ void col_body(const RpDev& d, int blk) {
  extern __shared__ float2 smem[];
  // the run state is checked once the loads are in flight (a stopped run's
  // kernels return before any store)
  float2* A = smem;
  float2* B = smem + fft_buf(N);
  float2* TW = smem + 2 * fft_buf(N);
  float* pa = (float*)(TW + N);  // phi hat columns a, b; Ey hat columns a, b
  float* pb = pa + N;
  float* ea = pb + N;
  float* eb = ea + N;
  const int ua = 2 * blk, ub = ua + 1;
  const size_t nx = (size_t)d.nbx;
  // Everything this thread reads from global memory in any phase is loaded
  // first (its rows v = threadIdx.x + r blockDim.x): the two columns, the
  // twiddles of the DCT post/pre steps and the spectral weights.
  constexpr int MAXC = (N + RP_BLOCK - 1) / RP_BLOCK;
  float2 colv[MAXC], tw[MAXC];
  float wyv[MAXC], wy2v[MAXC];
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int v = threadIdx.x + r * blockDim.x;
    if (v < N) {
      // columns ua, ub of the row-transformed grid t (row-major): one 8-byte
      // load per row (the workgroups sharing a cache line run on one XCD)
      colv[r] = *(const float2*)(d.t + v * nx + ua);
      tw[r] = d.tw2y[v];
      wyv[r] = d.wy[v];
      wy2v[r] = d.wy2[v];
    }
  }
  const float wx2u[2] = {d.wx2[ua], d.wx2[ub]};
  StageTw<N, RP_BLOCK> stw;
  stw.load(d.twy, threadIdx.x);
  if (run_phase(d.st) != 0) return;
  stw.store(TW, threadIdx.x);
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int v = threadIdx.x + r * blockDim.x;
    if (v < N) A[lpad(mk_pos(v, N))] = colv[r];
  }
  __syncthreads();
  const float2* Z = fft_n<N, -1>(A, B, TW);
  const float scale4 = (float)(4.0 / (double)d.nbx / (double)N);
#pragma unroll
  for (int rr = 0; rr < MAXC; rr++) {
    const int v = threadIdx.x + rr * blockDim.x;
    if (v >= N) continue;
    float xa, xb;
    dct2_post<N>(Z, v, tw[rr], xa, xb);
    const float wy = wyv[rr], wy2 = wy2v[rr];
    const float f[2] = {xa, xb};
    const int us[2] = {ua, ub};
    for (int r = 0; r < 2; r++) {
      const int u = us[r];
      float F = f[r];
      if (v == 0) F *= 0.5f;
      if (u == 0) F *= 0.5f;
      F = F * scale4;
      float ph = 0.0f, ey = 0.0f;
      if (!(u == 0 && v == 0)) {
        ph = F / (wx2u[r] + wy2);
        ey = ph * wy;
      }
      (r ? pb : pa)[v] = ph;
      (r ? eb : ea)[v] = ey;
    }
  }
  __syncthreads();
  // DCT-III of phi hat (columns a, b) -> y1
#pragma unroll
  for (int rr = 0; rr < MAXC; rr++) {
    const int k = threadIdx.x + rr * blockDim.x;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    const float ank = k ? pa[nk] : 0.0f, bnk = k ? pb[nk] : 0.0f;
    A[lpad(k)] = dct3_pre(pa[k], ank, pb[k], bnk, k, tw[rr]);
  }
  __syncthreads();
  Z = fft_n<N, +1>(A, B, TW);
  for (int m = threadIdx.x; m < N; m += blockDim.x)
    *(float2*)(d.y1 + m * nx + ua) = dct3_out<N>(Z, m);  // row-major, as the row pass reads it
  __syncthreads();
  // DST-III of Ey hat: b'[0] = a[0] (the A[N] slot), b'[j] = a[N-j]; sign (-1)^k
#pragma unroll
  for (int rr = 0; rr < MAXC; rr++) {
    const int k = threadIdx.x + rr * blockDim.x;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    const float ak = k ? ea[N - k] : ea[0];
    const float bk = k ? eb[N - k] : eb[0];
    const float ank = k ? (nk ? ea[N - nk] : ea[0]) : 0.0f;
    const float bnk = k ? (nk ? eb[N - nk] : eb[0]) : 0.0f;
    A[lpad(k)] = dct3_pre(ak, ank, bk, bnk, k, tw[rr]);
  }
  __syncthreads();
  Z = fft_n<N, +1>(A, B, TW);
  for (int m = threadIdx.x; m < N; m += blockDim.x) {
    const float2 z = dct3_out<N>(Z, m);
    const float sg = (m & 1) ? -1.0f : 1.0f;
    *(float2*)(d.y2 + m * nx + ua) = make_float2(sg * z.x, sg * z.y);
  }
}

// ---------------------------------------------------------------------------
// 5. inverse along x: Ex = DST-III_x(wx * Y1), Ey = DCT-III_x(Y2), phi optional
// ---------------------------------------------------------------------------
template <int N>
__device__ void rowinv_body(const RpDev& d, int blk, int want_phi) {
  extern __shared__ float2 smem[];
  // the run state is checked once the loads are in flight (a stopped run's
  // kernels return before any store)
  float2* A = smem;
  float2* B = smem + fft_buf(N);
  float2* TW = smem + 2 * fft_buf(N);
  float* ra = (float*)(TW + N);  // input rows a, b, then Ex rows a, b
  float* rb = ra + N;
  float* qa = rb + N;  // second input rows a, b (Y2, then Y1 for phi)
  float* qb = qa + N;
  const int ja = 2 * blk, jb = ja + 1;
  // the thread's loads of every phase first (its entries u = threadIdx.x +
  // r blockDim.x): the two rows of y1 and y2, wx and the DCT-III twiddles
  constexpr int MAXC = (N + RP_BLOCK - 1) / RP_BLOCK;
  float wv[MAXC], y1a[MAXC], y1b[MAXC], y2a[MAXC], y2b[MAXC];
  float2 tw[MAXC];
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int u = threadIdx.x + r * blockDim.x;
    if (u < N) {
      wv[r] = d.wx[u];
      y1a[r] = d.y1[(size_t)ja * N + u];
      y1b[r] = d.y1[(size_t)jb * N + u];
      y2a[r] = d.y2[(size_t)ja * N + u];
      y2b[r] = d.y2[(size_t)jb * N + u];
      tw[r] = d.tw2x[u];
    }
  }
  StageTw<N, RP_BLOCK> stw;
  stw.load(d.twx, threadIdx.x);
  if (run_phase(d.st) != 0) return;
  stw.store(TW, threadIdx.x);
  // Ex: DST-III of (wx[u] * Y1[j][u]); Ey: DCT-III of Y2 rows
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int u = threadIdx.x + r * blockDim.x;
    if (u < N) {
      ra[u] = wv[r] * y1a[r];
      rb[u] = wv[r] * y1b[r];
      qa[u] = y2a[r];
      qb[u] = y2b[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int k = threadIdx.x + r * blockDim.x;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    const float ak = k ? ra[N - k] : ra[0];
    const float bk = k ? rb[N - k] : rb[0];
    const float ank = k ? (nk ? ra[N - nk] : ra[0]) : 0.0f;
    const float bnk = k ? (nk ? rb[N - nk] : rb[0]) : 0.0f;
    A[lpad(k)] = dct3_pre(ak, ank, bk, bnk, k, tw[r]);
  }
  __syncthreads();
  const float2* Z = fft_n<N, +1>(A, B, TW);
  for (int m = threadIdx.x; m < N; m += blockDim.x) {
    const float2 z = dct3_out<N>(Z, m);
    const float sg = (m & 1) ? -1.0f : 1.0f;
    ra[m] = sg * z.x;  // Ex row a (ra no longer needed as input)
    rb[m] = sg * z.y;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int k = threadIdx.x + r * blockDim.x;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    A[lpad(k)] = dct3_pre(qa[k], k ? qa[nk] : 0.0f, qb[k], k ? qb[nk] : 0.0f, k, tw[r]);
  }
  __syncthreads();
  Z = fft_n<N, +1>(A, B, TW);
  for (int m = threadIdx.x; m < N; m += blockDim.x) {
    const float2 z = dct3_out<N>(Z, m);
    d.field[(size_t)ja * N + m] = make_float2(ra[m], z.x);
    d.field[(size_t)jb * N + m] = make_float2(rb[m], z.y);
  }
  if (!want_phi) return;
  __syncthreads();
  for (int u = threadIdx.x; u < N; u += blockDim.x) {
    qa[u] = d.y1[(size_t)ja * N + u];
    qb[u] = d.y1[(size_t)jb * N + u];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int k = threadIdx.x + r * blockDim.x;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    A[lpad(k)] = dct3_pre(qa[k], k ? qa[nk] : 0.0f, qb[k], k ? qb[nk] : 0.0f, k, tw[r]);
  }
  __syncthreads();
  Z = fft_n<N, +1>(A, B, TW);
  double vphi = 0.0;
  for (int m = threadIdx.x; m < N; m += blockDim.x) {
    const float2 z = dct3_out<N>(Z, m);
    d.phi[(size_t)ja * N + m] = z.x;
    d.phi[(size_t)jb * N + m] = z.y;
    if (d.part_phi)
      vphi += (double)(z.x * d.dbg_area[(size_t)ja * N + m]) + (double)(z.y * d.dbg_area[(size_t)jb * N + m]);
  }
  if (d.part_phi) {  // sumPhi (src/nesterovBase.cpp:1493-1505), verbose-3 trace only
    __shared__ double sh_phi[RP_BLOCK / 64];
    const double tot = block_sum_d(vphi, sh_phi);
    if (threadIdx.x == 0) d.part_phi[blk] = tot;
  }
}

// ---------------------------------------------------------------------------
// Wide forms of the column and inverse passes (single instance, RP_WIDE
// threads per workgroup): the two halves of the workgroup transform side by
// side.
//  * col4_body: four adjacent columns per workgroup (two column pairs, one
//    per half), read and written as 16-byte row segments, so a wavefront's
//    access touches half as many cache lines as two columns at a time.
//  * rowinv2_body: the two inverse transforms of a row pair (Ex and Ey) in
//    parallel, one per half.
// The arithmetic of every value is that of col_body / rowinv_body.
// ---------------------------------------------------------------------------
#define RP_WIDE 512

inline size_t rp_lds_col4(int n) { return sizeof(float2) * (4 * (size_t)(n + n / 16) + (size_t)n) + sizeof(float) * 8 * (size_t)n; }
inline size_t rp_lds_rinv2(int n) { return rp_lds_col4(n) - sizeof(float) * 4 * (size_t)n; }

template <int N>
__device__ void col4_body(const RpDev& d, int blk) {
  extern __shared__ float2 smem[];
  // the run state is checked once the loads are in flight (a stopped run's
  // kernels return before any store)
  const int h = threadIdx.x / RP_BLOCK, lt = threadIdx.x % RP_BLOCK;
  float2* A = smem + 2 * h * fft_buf(N);
  float2* B = A + fft_buf(N);
  float2* TW = smem + 4 * fft_buf(N);
  float* stage = (float*)(TW + N) + 4 * h * N;  // this half's phi hat columns a, b; Ey hat columns a, b
  float* pa = stage;
  float* pb = pa + N;
  float* ea = pb + N;
  float* eb = ea + N;
  const int u0 = 4 * blk, ua = u0 + 2 * h, ub = ua + 1;
  const size_t nx = (size_t)d.nbx;
  constexpr int MAXW = (N + RP_WIDE - 1) / RP_WIDE;   // rows per thread, whole workgroup
  constexpr int MAXC = (N + RP_BLOCK - 1) / RP_BLOCK; // rows per thread, one half
  float4 colv[MAXW];
  float2 tw[MAXC];
  float wyv[MAXC], wy2v[MAXC];
#pragma unroll
  for (int r = 0; r < MAXW; r++) {
    const int v = threadIdx.x + r * RP_WIDE;
    if (v < N) colv[r] = *(const float4*)(d.t + v * nx + u0);
  }
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int v = lt + r * RP_BLOCK;
    if (v < N) {
      tw[r] = d.tw2y[v];
      wyv[r] = d.wy[v];
      wy2v[r] = d.wy2[v];
    }
  }
  const float wx2u[2] = {d.wx2[ua], d.wx2[ub]};
  StageTw<N, RP_WIDE> stw;
  stw.load(d.twy, threadIdx.x);
  if (run_phase(d.st) != 0) return;
  stw.store(TW, threadIdx.x);
  float2* A0 = smem;
  float2* A1 = smem + 2 * fft_buf(N);
#pragma unroll
  for (int r = 0; r < MAXW; r++) {
    const int v = threadIdx.x + r * RP_WIDE;
    if (v < N) {
      const int p = lpad(mk_pos(v, N));
      A0[p] = make_float2(colv[r].x, colv[r].y);
      A1[p] = make_float2(colv[r].z, colv[r].w);
    }
  }
  __syncthreads();
  const float2* Z = fft_n<N, -1>(A, B, TW, lt, RP_BLOCK);
  const float scale4 = (float)(4.0 / (double)d.nbx / (double)N);
#pragma unroll
  for (int rr = 0; rr < MAXC; rr++) {
    const int v = lt + rr * RP_BLOCK;
    if (v >= N) continue;
    float xa, xb;
    dct2_post<N>(Z, v, tw[rr], xa, xb);
    const float wy = wyv[rr], wy2 = wy2v[rr];
    const float f[2] = {xa, xb};
    const int us[2] = {ua, ub};
    for (int r = 0; r < 2; r++) {
      const int u = us[r];
      float F = f[r];
      if (v == 0) F *= 0.5f;
      if (u == 0) F *= 0.5f;
      F = F * scale4;
      float ph = 0.0f, ey = 0.0f;
      if (!(u == 0 && v == 0)) {
        ph = F / (wx2u[r] + wy2);
        ey = ph * wy;
      }
      (r ? pb : pa)[v] = ph;
      (r ? eb : ea)[v] = ey;
    }
  }
  __syncthreads();
  // DCT-III of phi hat -> y1
#pragma unroll
  for (int rr = 0; rr < MAXC; rr++) {
    const int k = lt + rr * RP_BLOCK;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    const float ank = k ? pa[nk] : 0.0f, bnk = k ? pb[nk] : 0.0f;
    A[lpad(k)] = dct3_pre(pa[k], ank, pb[k], bnk, k, tw[rr]);
  }
  __syncthreads();
  Z = fft_n<N, +1>(A, B, TW, lt, RP_BLOCK);
  // both halves' results sit in the same array of their pair of buffers
  const float2* Z0 = Z - 2 * h * fft_buf(N);
  const float2* Z1 = Z0 + 2 * fft_buf(N);
  for (int m = threadIdx.x; m < N; m += RP_WIDE) {
    const float2 z0 = dct3_out<N>(Z0, m), z1 = dct3_out<N>(Z1, m);
    *(float4*)(d.y1 + m * nx + u0) = make_float4(z0.x, z0.y, z1.x, z1.y);
  }
  __syncthreads();
  // DST-III of Ey hat: b'[0] = a[0] (the A[N] slot), b'[j] = a[N-j]; sign (-1)^k
#pragma unroll
  for (int rr = 0; rr < MAXC; rr++) {
    const int k = lt + rr * RP_BLOCK;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    const float ak = k ? ea[N - k] : ea[0];
    const float bk = k ? eb[N - k] : eb[0];
    const float ank = k ? (nk ? ea[N - nk] : ea[0]) : 0.0f;
    const float bnk = k ? (nk ? eb[N - nk] : eb[0]) : 0.0f;
    A[lpad(k)] = dct3_pre(ak, ank, bk, bnk, k, tw[rr]);
  }
  __syncthreads();
  Z = fft_n<N, +1>(A, B, TW, lt, RP_BLOCK);
  const float2* W0 = Z - 2 * h * fft_buf(N);
  const float2* W1 = W0 + 2 * fft_buf(N);
  for (int m = threadIdx.x; m < N; m += RP_WIDE) {
    const float2 z0 = dct3_out<N>(W0, m), z1 = dct3_out<N>(W1, m);
    const float sg = (m & 1) ? -1.0f : 1.0f;
    *(float4*)(d.y2 + m * nx + u0) = make_float4(sg * z0.x, sg * z0.y, sg * z1.x, sg * z1.y);
  }
}

template <int N>
__device__ void rowinv2_body(const RpDev& d, int blk, int want_phi) {
  extern __shared__ float2 smem[];
  // the run state is checked once the loads are in flight (a stopped run's
  // kernels return before any store)
  const int h = threadIdx.x / RP_BLOCK, lt = threadIdx.x % RP_BLOCK;
  float2* A = smem + 2 * h * fft_buf(N);
  float2* B = A + fft_buf(N);
  float2* TW = smem + 4 * fft_buf(N);
  float* ra = (float*)(TW + N);  // Ex input rows a, b (wx * Y1)
  float* rb = ra + N;
  float* qa = rb + N;  // Ey input rows a, b (Y2)
  float* qb = qa + N;
  const int ja = 2 * blk, jb = ja + 1;
  constexpr int MAXW = (N + RP_WIDE - 1) / RP_WIDE;
  constexpr int MAXC = (N + RP_BLOCK - 1) / RP_BLOCK;
  float wv[MAXW], y1a[MAXW], y1b[MAXW], y2a[MAXW], y2b[MAXW];
  float2 tw[MAXC];
#pragma unroll
  for (int r = 0; r < MAXW; r++) {
    const int u = threadIdx.x + r * RP_WIDE;
    if (u < N) {
      wv[r] = d.wx[u];
      y1a[r] = d.y1[(size_t)ja * N + u];
      y1b[r] = d.y1[(size_t)jb * N + u];
      y2a[r] = d.y2[(size_t)ja * N + u];
      y2b[r] = d.y2[(size_t)jb * N + u];
    }
  }
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int k = lt + r * RP_BLOCK;
    if (k < N) tw[r] = d.tw2x[k];
  }
  StageTw<N, RP_WIDE> stw;
  stw.load(d.twx, threadIdx.x);
  if (run_phase(d.st) != 0) return;
  stw.store(TW, threadIdx.x);
#pragma unroll
  for (int r = 0; r < MAXW; r++) {
    const int u = threadIdx.x + r * RP_WIDE;
    if (u < N) {
      ra[u] = wv[r] * y1a[r];
      rb[u] = wv[r] * y1b[r];
      qa[u] = y2a[r];
      qb[u] = y2b[r];
    }
  }
  __syncthreads();
  // half 0: DST-III of wx * Y1 (Ex); half 1: DCT-III of Y2 (Ey)
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int k = lt + r * RP_BLOCK;
    if (k >= N) continue;
    const int nk = (N - k) & (N - 1);
    if (h == 0) {
      const float ak = k ? ra[N - k] : ra[0];
      const float bk = k ? rb[N - k] : rb[0];
      const float ank = k ? (nk ? ra[N - nk] : ra[0]) : 0.0f;
      const float bnk = k ? (nk ? rb[N - nk] : rb[0]) : 0.0f;
      A[lpad(k)] = dct3_pre(ak, ank, bk, bnk, k, tw[r]);
    } else {
      A[lpad(k)] = dct3_pre(qa[k], k ? qa[nk] : 0.0f, qb[k], k ? qb[nk] : 0.0f, k, tw[r]);
    }
  }
  __syncthreads();
  const float2* Z = fft_n<N, +1>(A, B, TW, lt, RP_BLOCK);
  const float2* Z0 = Z - 2 * h * fft_buf(N);
  const float2* Z1 = Z0 + 2 * fft_buf(N);
  for (int m = threadIdx.x; m < N; m += RP_WIDE) {
    const float2 zx = dct3_out<N>(Z0, m), zy = dct3_out<N>(Z1, m);
    const float sg = (m & 1) ? -1.0f : 1.0f;
    d.field[(size_t)ja * N + m] = make_float2(sg * zx.x, zy.x);
    d.field[(size_t)jb * N + m] = make_float2(sg * zx.y, zy.y);
  }
  if (!want_phi) return;
  // phi: DCT-III of Y1 rows (half 0; half 1 only meets the barriers)
  __syncthreads();
  for (int u = threadIdx.x; u < N; u += RP_WIDE) {
    qa[u] = d.y1[(size_t)ja * N + u];
    qb[u] = d.y1[(size_t)jb * N + u];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MAXC; r++) {
    const int k = lt + r * RP_BLOCK;
    if (h != 0 || k >= N) continue;
    const int nk = (N - k) & (N - 1);
    A[lpad(k)] = dct3_pre(qa[k], k ? qa[nk] : 0.0f, qb[k], k ? qb[nk] : 0.0f, k, tw[r]);
  }
  __syncthreads();
  Z = fft_n<N, +1>(A, B, TW, lt, RP_BLOCK);
  double vphi = 0.0;
  if (h == 0)
    for (int m = lt; m < N; m += RP_BLOCK) {
      const float2 z = dct3_out<N>(Z, m);
      d.phi[(size_t)ja * N + m] = z.x;
      d.phi[(size_t)jb * N + m] = z.y;
      if (d.part_phi)
        vphi += (double)(z.x * d.dbg_area[(size_t)ja * N + m]) + (double)(z.y * d.dbg_area[(size_t)jb * N + m]);
    }
  if (d.part_phi) {  // sumPhi (src/nesterovBase.cpp:1493-1505), verbose-3 trace only
    __shared__ double sh_phi[RP_WIDE / 64];
    const double tot = block_sum_d(vphi, sh_phi);
    if (threadIdx.x == 0) d.part_phi[blk] = tot;
  }
}

// Calls BODY<n>(args) for the transform length n of the grid side at hand.
#define RP_FFT_DISPATCH(n, BODY, ...)        \
  switch (n) {                               \
    case 4: BODY<4>(__VA_ARGS__); break;     \
    case 8: BODY<8>(__VA_ARGS__); break;     \
    case 16: BODY<16>(__VA_ARGS__); break;   \
    case 32: BODY<32>(__VA_ARGS__); break;   \
    case 64: BODY<64>(__VA_ARGS__); break;   \
    case 128: BODY<128>(__VA_ARGS__); break; \
    case 256: BODY<256>(__VA_ARGS__); break; \
    case 512: BODY<512>(__VA_ARGS__); break; \
    case 1024: BODY<1024>(__VA_ARGS__); break; \
    case 2048: BODY<2048>(__VA_ARGS__); break; \
    default: break;                          \
  }

// ---------------------------------------------------------------------------
// 6. WA exponent sums per net (src/nesterovBase.cpp:1304-1360) + HPWL
// ---------------------------------------------------------------------------

// The four WA exponentials of a pin (src/nesterovBase.cpp:1316-1356): the
// min-x, max-x, min-y and max-y terms against the net's pin bbox, each
// fastExp(exponent) where the exponent exceeds the force bar and -1 (never a
// fastExp value, which is >= 0) where it does not.
__device__ __forceinline__ float4 wa_exps(int2 p, int4 box, float cx, float cy, float bar) {
  float e0 = (float)(box.x - p.x) * cx;
  float e1 = (float)(p.x - box.z) * cx;
  float e2 = (float)(box.y - p.y) * cy;
  float e3 = (float)(p.y - box.w) * cy;
  return make_float4(e0 > bar ? fast_exp(e0) : -1.0f, e1 > bar ? fast_exp(e1) : -1.0f,
                     e2 > bar ? fast_exp(e2) : -1.0f, e3 > bar ? fast_exp(e3) : -1.0f);
}

// The eight per-net sums, accumulated in pin order.
struct WaSums {
  float s0, s1, s2, s3, s4, s5, s6, s7;
};

__device__ __forceinline__ void wa_accumulate(WaSums& w, int2 p, float4 v) {
  if (v.x >= 0.0f) {
    w.s0 += v.x;
    w.s1 += (float)p.x * v.x;
  }
  if (v.y >= 0.0f) {
    w.s2 += v.y;
    w.s3 += (float)p.x * v.y;
  }
  if (v.z >= 0.0f) {
    w.s4 += v.z;
    w.s5 += (float)p.y * v.z;
  }
  if (v.w >= 0.0f) {
    w.s6 += v.w;
    w.s7 += (float)p.y * v.w;
  }
}

// getWireLengthGradientPinWA for one pin and one axis (src/nesterovBase.cpp:1383-1440):
// gMin - gMax from the pin's exponentials (negative = inactive) and the net sums.
__device__ __forceinline__ float wl_axis(int p, float coef, float ex0, float ex1, float S0, float XS0, float S1,
                                         float XS1) {
  float gmin = 0.0f, gmax = 0.0f;
  if (ex0 >= 0.0f)
    gmin = (float)(((double)S0 * ((double)ex0 * (1.0 - (double)(coef * (float)p))) + (double)(coef * ex0 * XS0)) /
                   (double)(S0 * S0));
  if (ex1 >= 0.0f)
    gmax = (float)(((double)S1 * ((double)ex1 * (1.0 + (double)(coef * (float)p))) - (double)(coef * ex1 * XS1)) /
                   (double)(S1 * S1));
  return gmin - gmax;
}

__device__ __forceinline__ float2 wa_pin_grad(int2 p, float4 v, const float4& sa, const float4& sb, float cx,
                                              float cy) {
  return make_float2(wl_axis(p.x, cx, v.x, v.y, sa.x, sa.y, sa.z, sa.w),
                     wl_axis(p.y, cy, v.z, v.w, sb.x, sb.y, sb.z, sb.w));
}

// Weighted-average wirelength per workgroup of nets (updateWireLengthForceWA
// and getWireLengthGradientPinWA, src/nesterovBase.cpp:1304-1440), pin
// parallel.  The workgroup takes the nets d.wa_rng[blk] (consecutive, at
// most RP_WA_NETS nets and CAP pins, or one larger net alone; the packing
// for each of the RP_WA_VARIANTS capacities is built at set-up):
//   1. pin centres (cell centre + offset, pin_centre) into LDS;
//   2. one lane per net: pin bbox, the pin -> net map, HPWL term;
//   3. one lane per pin: the four exponentials;
//   4. one lane per net: the eight sums in pin order (the reference's float
//      summation order, so the sums are bit-identical);
//   5. one lane per pin: the per-pin gradient gMin - gMax, stored in pin
//      order (coalesced; the cell gradient kernel gathers it).
// The net bbox and sums are stored only when store_nets is set (stage-level
// checks).
template <int CAP>
struct WaLds {
  float4 ex[CAP];
  union {  // the net boxes are dead once the exponentials are written, the sums come after
    float4 sum[2 * RP_WA_NETS];
    int4 box[RP_WA_NETS];
  };
  int2 pos[CAP];
  long long red[RP_BLOCK / 64];
  uint8_t net[CAP];  // net index within the workgroup (< RP_WA_NETS = 256)
};
template <int CAP>
__device__ void wa_body(const RpDev& d, int blk, int store_nets, WaLds<CAP>& W) {
  int2* s_pos = W.pos;
  float4* s_ex = W.ex;
  uint8_t* s_net = W.net;
  int4* s_box = W.box;
  float4* s_sum = W.sum;
  long long* sh_red = W.red;
  const RpState* st = d.st;
  if (st->phase != 0) return;
  const float cx = st->wl_coef_x, cy = st->wl_coef_y, bar = d.bar;
  const int4 rng = d.wa_rng[blk];
  const int p0 = rng.z, p1 = rng.w;
  const int npw = p1 - p0;
  const int tid = threadIdx.x;
  long long h = 0;
  if (npw <= CAP) {
    // Every phase issues its loads before it consumes them (the pins of a
    // thread are tid + RP_BLOCK r); the per-net loops read four pins ahead.
    // The arithmetic and its order are those of one pin at a time.
    constexpr int PER = (CAP + RP_BLOCK - 1) / RP_BLOCK;
    const int e = rng.x + tid;
    const bool owner = e < rng.y;
    int s = 0, t = 0;
    if (owner) {  // issued beside the pin loads
      s = d.nstart[e] - p0;
      t = d.nstart[e + 1] - p0;
    }
    {
      // pin centres (pin_centre): cell ids and offsets of all of a thread's
      // pins first, then the cell centre gathers
      int pc[PER];
      int2 lp[PER];
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int j = tid + r * RP_BLOCK;
        if (j < npw) {
          pc[r] = d.pcell[p0 + j];
          lp[r] = d.ppos[p0 + j];
        }
      }
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int j = tid + r * RP_BLOCK;
        if (j < npw && pc[r] >= 0) {
          const int2 icc = d.ic[pc[r]];
          lp[r] = make_int2(icc.x + lp[r].x, icc.y + lp[r].y);
        }
      }
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int j = tid + r * RP_BLOCK;
        if (j < npw) s_pos[j] = lp[r];
      }
    }
    __syncthreads();
    int4 box = make_int4(INT_MAX, INT_MAX, INT_MIN, INT_MIN);
    if (owner) {
      int q = s;
      for (; q + 4 <= t; q += 4) {
        int2 p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) p[u] = s_pos[q + u];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          box.x = min(p[u].x, box.x);
          box.y = min(p[u].y, box.y);
          box.z = max(p[u].x, box.z);
          box.w = max(p[u].y, box.w);
          s_net[q + u] = (uint8_t)tid;
        }
      }
      for (; q < t; q++) {
        int2 p = s_pos[q];
        box.x = min(p.x, box.x);
        box.y = min(p.y, box.y);
        box.z = max(p.x, box.z);
        box.w = max(p.y, box.w);
        s_net[q] = (uint8_t)tid;
      }
      s_box[tid] = box;
      if (t > s) h = (long long)((box.z - box.x) + (box.w - box.y));
    }
    __syncthreads();
    {
      int net[PER];
      int2 pos[PER];
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int j = tid + r * RP_BLOCK;
        if (j < npw) {
          net[r] = s_net[j];
          pos[r] = s_pos[j];
        }
      }
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int j = tid + r * RP_BLOCK;
        if (j < npw) s_ex[j] = wa_exps(pos[r], s_box[net[r]], cx, cy, bar);
      }
    }
    __syncthreads();
    if (owner) {
      WaSums w = {0, 0, 0, 0, 0, 0, 0, 0};
      int q = s;
      for (; q + 4 <= t; q += 4) {
        int2 p[4];
        float4 x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          p[u] = s_pos[q + u];
          x[u] = s_ex[q + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) wa_accumulate(w, p[u], x[u]);
      }
      for (; q < t; q++) wa_accumulate(w, s_pos[q], s_ex[q]);
      float4 sa = make_float4(w.s0, w.s1, w.s2, w.s3), sb = make_float4(w.s4, w.s5, w.s6, w.s7);
      s_sum[2 * tid] = sa;
      s_sum[2 * tid + 1] = sb;
      if (store_nets) {
        d.nbox[e] = box;
        d.nsum[2 * (size_t)e] = sa;
        d.nsum[2 * (size_t)e + 1] = sb;
      }
    }
    __syncthreads();
    {
      int net[PER];
      int2 pos[PER];
      float4 ex[PER];
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int j = tid + r * RP_BLOCK;
        if (j < npw) {
          net[r] = s_net[j];
          pos[r] = s_pos[j];
          ex[r] = s_ex[j];
        }
      }
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const int j = tid + r * RP_BLOCK;
        if (j < npw) {
          const int L = net[r];
          d.pgrad[p0 + j] = wa_pin_grad(pos[r], ex[r], s_sum[2 * L], s_sum[2 * L + 1], cx, cy);
        }
      }
    }
  } else {
    // one net with more than CAP pins: bbox by a block min/max, the
    // sums chunk by chunk (exponentials in parallel, summed in pin order by
    // lane 0), then the per-pin gradients with the exponentials recomputed
    // (same arguments, same values)
    const int e = rng.x;
    int4 box = make_int4(INT_MAX, INT_MAX, INT_MIN, INT_MIN);
    for (int q = tid; q < npw; q += blockDim.x) {
      int2 p = pin_centre(d, p0 + q);
      box.x = min(p.x, box.x);
      box.y = min(p.y, box.y);
      box.z = max(p.x, box.z);
      box.w = max(p.y, box.w);
    }
    for (int o = 32; o > 0; o >>= 1) {
      box.x = min(box.x, __shfl_xor(box.x, o, 64));
      box.y = min(box.y, __shfl_xor(box.y, o, 64));
      box.z = max(box.z, __shfl_xor(box.z, o, 64));
      box.w = max(box.w, __shfl_xor(box.w, o, 64));
    }
    if ((tid & 63) == 0) s_box[tid >> 6] = box;
    __syncthreads();
    box = s_box[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
      int4 b = s_box[k];
      box.x = min(b.x, box.x);
      box.y = min(b.y, box.y);
      box.z = max(b.z, box.z);
      box.w = max(b.w, box.w);
    }
    WaSums w = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c0 = 0; c0 < npw; c0 += CAP) {
      const int cnt = min(CAP, npw - c0);
      __syncthreads();  // previous chunk consumed (and s_box read by everyone)
      for (int j = tid; j < cnt; j += blockDim.x) {
        int2 p = pin_centre(d, p0 + c0 + j);
        s_pos[j] = p;
        s_ex[j] = wa_exps(p, box, cx, cy, bar);
      }
      __syncthreads();
      if (tid == 0)
        for (int j = 0; j < cnt; j++) wa_accumulate(w, s_pos[j], s_ex[j]);
    }
    if (tid == 0) {
      s_sum[0] = make_float4(w.s0, w.s1, w.s2, w.s3);
      s_sum[1] = make_float4(w.s4, w.s5, w.s6, w.s7);
      h = (long long)((box.z - box.x) + (box.w - box.y));
      if (store_nets) {
        d.nbox[e] = box;
        d.nsum[2 * (size_t)e] = s_sum[0];
        d.nsum[2 * (size_t)e + 1] = s_sum[1];
      }
    }
    __syncthreads();
    const float4 sa = s_sum[0], sb = s_sum[1];
    for (int q = tid; q < npw; q += blockDim.x) {
      int2 p = pin_centre(d, p0 + q);
      d.pgrad[p0 + q] = wa_pin_grad(p, wa_exps(p, box, cx, cy, bar), sa, sb, cx, cy);
    }
  }
  long long tot = block_sum_ll(h, sh_red);
  if (threadIdx.x == 0) d.part_net[blk] = tot;
}

// wa_body with its own LDS (the wirelength kernels of one kind)
template <int CAP>
__device__ void wa_body(const RpDev& d, int blk, int store_nets) {
  __shared__ WaLds<CAP> W;
  wa_body<CAP>(d, blk, store_nets, W);
}

// ---------------------------------------------------------------------------
// 7. gradients per gcell (src/nesterovPlace.cpp:206-273)
// The controller runs as its own one-workgroup kernel after this one: a
// last-block-done fusion needs an agent-scope release per block, which on
// MI355X writes back the XCD's L2 and costs far more than one launch.
__device__ void grad_body(const RpDev& d, int blk) {
  __shared__ double sh_red[RP_P_COUNT][RP_BLOCK / 64];
  const RpState* st = d.st;
  if (st->phase != 0) return;
  int c = blk * blockDim.x + threadIdx.x;
  double v_wl = 0.0, v_dens = 0.0, v_grad = 0.0, v_dc = 0.0, v_dg = 0.0, v_gn = 0.0;
  if (c < d.nc) {
    float lam = st->density_penalty;
    int dst = st->dst, ref = st->ring.cur;
    // Every load that needs only the cell index is issued first.  The pin
    // and bin gathers below stay one small batch or one bin at a time:
    // prefetching a 3 x 3 field box for every cell (up to 9 gathers instead
    // of the ~6 a box covers) measured slower, the gathers' issue rate and
    // not their latency being what this kernel waits on.
    const int2 icc = d.ic[c];
    const int2 h = d.dhalf[c];
    const float sc = d.scale[c];
    const float np = d.npin[c], ar = d.area[c];
    const int k0 = d.cp_start[c], k1 = d.cp_start[c + 1];
    // the distance terms' operands, loaded before the gradient store below
    // (which the compiler cannot tell apart from them)
    const float2 a = d.slp[(size_t)ref * d.nc + c];
    const float2 b = d.slp[(size_t)dst * d.nc + c];
    const float2 ga = d.sumg[(size_t)ref * d.nc + c];
    const int dlx = icc.x - h.x, dux = icc.x + h.x, dly = icc.y - h.y, duy = icc.y + h.y;
    int x0, x1, y0, y1;
    idx_range(dlx, dux, d.core_lx, d.bsx, d.nbx, x0, x1);
    idx_range(dly, duy, d.core_ly, d.bsy, d.nby, y0, y1);
    // wire length: per-pin WA gradients (written by k_wa), summed in the
    // cell's pin order (getWireLengthGradientWA, src/nesterovBase.cpp:1363-1375)
    float gx = 0.0f, gy = 0.0f;
    for (int k = k0; k < k1; k += 4) {
      int qq[4];
      float2 pg[4];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (k + j < k1) qq[j] = d.cp_pin[k + j];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (k + j < k1) pg[j] = d.pgrad[qq[j]];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (k + j < k1) {
          gx += pg[j].x;
          gy += pg[j].y;
        }
    }
    // field gather (getDensityGradient, src/nesterovBase.cpp:1458-1478)
    float fx = 0.0f, fy = 0.0f;
    for (int i = x0; i < x1; i++) {
      int blx, bux;
      bin_x(d, i, blx, bux);
      int w = min(bux, dux) - max(blx, dlx);
      for (int j = y0; j < y1; j++) {
        int bly, buy;
        bin_y(d, j, bly, buy);
        int hh = min(buy, duy) - max(bly, dly);
        float ov = (w <= 0 || hh <= 0) ? 0.0f : (float)w * (float)hh;
        ov = ov * sc;
        float2 f = d.field[(size_t)j * d.nbx + i];
        fx += ov * f.x;
        fy += ov * f.y;
      }
    }
    // sum gradient and preconditioner
    float sx = gx + lam * fx;
    float sy = gy + lam * fy;
    float ppx = np + lam * ar;
    float ppy = np + lam * ar;
    if (ppx <= d.min_precond) ppx = d.min_precond;
    if (ppy <= d.min_precond) ppy = d.min_precond;
    sx /= ppx;
    sy /= ppy;
    d.sumg[(size_t)dst * d.nc + c] = make_float2(sx, sy);
    if (d.dbg_wl) d.dbg_wl[c] = make_float2(gx, gy);
    if (d.dbg_dens) d.dbg_dens[c] = make_float2(fx, fy);
    v_wl = (double)fabsf(gx) + (double)fabsf(gy);
    v_dens = (double)fabsf(fx) + (double)fabsf(fy);
    v_grad = (double)(fabsf(sx) + fabsf(sy));
    // getDistance terms (src/nesterovPlace.cpp:598-607)
    float ddx = a.x - b.x, ddy = a.y - b.y;
    float gdx = ga.x - sx, gdy = ga.y - sy;
    v_dc = (double)(ddx * ddx) + (double)(ddy * ddy);
    v_dg = (double)(gdx * gdx) + (double)(gdy * gdy);
    v_gn = (double)(sx * sx + sy * sy);
  }
  if (d.part_gn) {  // getSecondNorm of the new gradient (src/nesterovPlace.cpp:609-616), trace only
    __shared__ double sh_gn[RP_BLOCK / 64];
    const double tot = block_sum_d(v_gn, sh_gn);
    if (threadIdx.x == 0) d.part_gn[blk] = tot;
  }
  // the five block sums in one pass: per value the same wavefront tree and
  // wavefront order as block_sum_d
  double v[RP_P_COUNT];
  v[RP_P_WL] = v_wl;
  v[RP_P_DENS] = v_dens;
  v[RP_P_GRAD] = v_grad;
  v[RP_P_DCOORD] = v_dc;
  v[RP_P_DGRAD] = v_dg;
#pragma unroll
  for (int m = 0; m < RP_P_COUNT; m++) v[m] = wave_sum_d(v[m]);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int m = 0; m < RP_P_COUNT; m++) sh_red[m][wv] = v[m];
  __syncthreads();
  if (threadIdx.x < RP_P_COUNT) {
    const int m = threadIdx.x;
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh_red[m][k];
    d.part_grad[(size_t)blk * RP_P_COUNT + m] = r;
  }
}

// ---------------------------------------------------------------------------
// 8. controller
// ---------------------------------------------------------------------------
// Fixed-order reduction of `n` partials with stride `stride` (1 workgroup).
// Fixed-order reductions of the per-block partials (one workgroup).
// Each thread sums a fixed strided subset in index order; the loads of
// RP_RED_BATCH iterations are issued before their additions (the sum order
// is unchanged) so the loop does not wait on one load at a time.
#define RP_RED_BATCH 8
__device__ double reduce_parts_d(const double* p, int n, int stride, int off, double* sh) {
  double v = 0.0;
  for (int k0 = threadIdx.x; k0 < n; k0 += RP_RED_BATCH * blockDim.x) {
    double x[RP_RED_BATCH];
#pragma unroll
    for (int u = 0; u < RP_RED_BATCH; u++) {
      int k = k0 + u * blockDim.x;
      x[u] = k < n ? p[(size_t)k * stride + off] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < RP_RED_BATCH; u++)
      if (k0 + u * (int)blockDim.x < n) v += x[u];
  }
  return block_sum_d(v, sh);
}

__device__ long long reduce_parts_ll(const long long* p, int n, long long* sh) {
  long long v = 0;
  for (int k0 = threadIdx.x; k0 < n; k0 += RP_RED_BATCH * blockDim.x) {
    long long x[RP_RED_BATCH];
#pragma unroll
    for (int u = 0; u < RP_RED_BATCH; u++) {
      int k = k0 + u * blockDim.x;
      x[u] = k < n ? p[k] : 0;
    }
#pragma unroll
    for (int u = 0; u < RP_RED_BATCH; u++) v += x[u];
  }
  return block_sum_ll(v, sh);
}

struct Totals {
  double wl, dens, grad, dc, dg, ovf;
  long long hpwl;
};

// Totals of the gradient, bin and net partials; valid in thread 0.
__device__ void reduce_all(const RpDev& d, Totals& t) {
  __shared__ long long shl[RP_CTRL_BLOCK / 64];
  // Each thread sums a fixed strided subset of the gradient, bin (overflow)
  // and net (HPWL) partials, batch by batch in index order; the first batch
  // of all three arrays is loaded before any of them is added.
  double v[RP_P_COUNT] = {0, 0, 0, 0, 0};
  double vb = 0.0;
  long long vh = 0;
  const int ng = d.npart_grad;
  const int nbp = d.npart_bin, nnp = d.nwa;  // one HPWL partial per wirelength workgroup
  constexpr int GB = 4;  // gradient iterations whose loads are issued together
  const int T = blockDim.x, k0 = threadIdx.x;
  // the net partials (one per wirelength workgroup: 8252 at 1M cells) in
  // batches of 16 per thread, so the usual sizes take one load round
  constexpr int NB = 16;
  double xg[GB][RP_P_COUNT], xb[RP_RED_BATCH];
  long long xn[NB];
#pragma unroll
  for (int u = 0; u < GB; u++) {
    const int k = k0 + u * T;
    const double* q = d.part_grad + (size_t)k * RP_P_COUNT;
#pragma unroll
    for (int j = 0; j < RP_P_COUNT; j++) xg[u][j] = k < ng ? q[j] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < RP_RED_BATCH; u++) {
    const int k = k0 + u * T;
    xb[u] = k < nbp ? d.part_bin[k] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int k = k0 + u * T;
    xn[u] = k < nnp ? d.part_net[k] : 0;
  }
  for (int kb = k0;;) {  // gradient partials
#pragma unroll
    for (int u = 0; u < GB; u++)
      if (kb + u * T < ng)
#pragma unroll
        for (int j = 0; j < RP_P_COUNT; j++) v[j] += xg[u][j];
    kb += GB * T;
    if (kb >= ng) break;
#pragma unroll
    for (int u = 0; u < GB; u++) {
      const int k = kb + u * T;
      const double* q = d.part_grad + (size_t)k * RP_P_COUNT;
#pragma unroll
      for (int j = 0; j < RP_P_COUNT; j++) xg[u][j] = k < ng ? q[j] : 0.0;
    }
  }
  for (int kb = k0;;) {  // bin partials
#pragma unroll
    for (int u = 0; u < RP_RED_BATCH; u++)
      if (kb + u * T < nbp) vb += xb[u];
    kb += RP_RED_BATCH * T;
    if (kb >= nbp) break;
#pragma unroll
    for (int u = 0; u < RP_RED_BATCH; u++) {
      const int k = kb + u * T;
      xb[u] = k < nbp ? d.part_bin[k] : 0.0;
    }
  }
  for (int kb = k0;;) {  // net partials (integers: any order)
#pragma unroll
    for (int u = 0; u < NB; u++) vh += xn[u];
    kb += NB * T;
    if (kb >= nnp) break;
#pragma unroll
    for (int u = 0; u < NB; u++) {
      const int k = kb + u * T;
      xn[u] = k < nnp ? d.part_net[k] : 0;
    }
  }
  // all seven totals in one pass: wavefront trees, then the wavefronts in
  // order (the same arithmetic as one block_sum per value)
  double w[RP_P_COUNT + 1];
#pragma unroll
  for (int j = 0; j < RP_P_COUNT; j++) w[j] = wave_sum_d(v[j]);
  w[RP_P_COUNT] = wave_sum_d(vb);
  vh = wave_sum_ll(vh);
  __shared__ double shw[RP_CTRL_BLOCK / 64][RP_P_COUNT + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j <= RP_P_COUNT; j++) shw[wv][j] = w[j];
    shl[wv] = vh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[RP_P_COUNT + 1] = {0, 0, 0, 0, 0, 0};
    long long h = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
#pragma unroll
      for (int j = 0; j <= RP_P_COUNT; j++) r[j] += shw[k][j];
      h += shl[k];
    }
    t.wl = r[RP_P_WL];
    t.dens = r[RP_P_DENS];
    t.grad = r[RP_P_GRAD];
    t.dc = r[RP_P_DCOORD];
    t.dg = r[RP_P_DGRAD];
    t.ovf = r[RP_P_COUNT];
    t.hpwl = h;
  }
  __syncthreads();
}

// Trace totals of one evaluation (all threads of the controller take part):
// the squared norm of the new gradient and sum phi * area.
struct TraceTotals {
  double gn, phi;
};
__device__ void reduce_trace(const RpDev& d, TraceTotals& tt) {
  __shared__ double sh[RP_CTRL_BLOCK / 64];
  tt.gn = reduce_parts_d(d.part_gn, d.st->num_partials_grad, 1, 0, sh);
  tt.phi = d.part_phi ? reduce_parts_d(d.part_phi, d.nby / 2, 1, 0, sh) : 0.0;
}

// updateWireLengthCoef, src/nesterovPlace.cpp:474-490
__device__ void update_wl_coef(RpState* s, float overflow) {
  float c;
  if (overflow > 1.0)
    c = 0.1f;
  else if (overflow < 0.1)
    c = 10.0f;
  else
    c = (float)(1.0 / pow(10.0, ((double)overflow - 0.1) * 20 / 9.0 - 1.0));
  s->wl_coef_x = c * s->base_wl_coef;
  s->wl_coef_y = c * s->base_wl_coef;
}

__device__ float overflow_of(const RpState* s, long long area) { return (float)area / s->area_denom; }

// getStepLength with FP64 sums: float(sqrt(sum / (2N))) per distance
__device__ float step_of(const RpState* s, double dc, double dg) {
  double n2 = 2.0 * (double)s->num_cells;
  float cd = (float)sqrt(dc / n2);
  float gd = (float)sqrt(dg / n2);
  return cd / gd;
}

// The verbose-3 record of the evaluation just reduced (thread 0).
__device__ void write_eval_rec(const RpDev& d, const RpState* s, const Totals& t, const TraceTotals& tt, int iter,
                               int trial, int diverged, float new_step, float overflow) {
  const int idx = s->evals - 1;
  if (idx < 0 || idx >= d.evrec_cap) return;
  const double n2 = 2.0 * (double)s->num_cells;
  RpEvalRec e;
  e.iter = iter;
  e.trial = trial;
  e.diverged = diverged;
  e.density_penalty = s->density_penalty;
  e.wl_coef = s->wl_coef_x;
  e.wl_grad_sum = (float)t.wl;
  e.density_grad_sum = (float)t.dens;
  e.grad_sum = (float)t.grad;
  e.coordinate_distance = (float)sqrt(t.dc / n2);
  e.gradient_distance = (float)sqrt(t.dg / n2);
  e.new_step_length = new_step;
  e.gradient_norm = (float)sqrt(tt.gn / n2);
  e.sum_phi = (float)tt.phi;
  e.overflow = overflow;
  e.hpwl = t.hpwl;
  d.evrec[idx] = e;
}

__device__ bool bad_sum(double v) {
  float f = (float)v;
  return isnan(f) || isinf(f);
}

__device__ void begin_iteration(RpState* s) {
  // src/nesterovPlace.cpp:318-326
  float prevA = s->cur_a;
  s->prev_a = prevA;
  s->cur_a = (float)((1.0 + sqrt(4.0 * (double)prevA * (double)prevA + 1.0)) * 0.5);
  s->coeff = (float)(((double)prevA - 1.0) / (double)s->cur_a);
  s->bt = 0;
  s->mode = 0;
  s->dst = s->ring.next;
}

// stage: 0 = after init eval A's Poisson solve (before WA)
//        1 = after init eval A's gradients
//        2 = after init eval B's gradients
__global__ void __launch_bounds__(RP_CTRL_BLOCK) k_control_init(RpDev d, int stage) {
  Totals t;
  reduce_all(d, t);
  TraceTotals tt = {0.0, 0.0};
  if (d.evrec && stage > 0) reduce_trace(d, tt);
  if (threadIdx.x != 0) return;
  RpState* s = d.st;
  if (s->phase != 0) return;
  long long ovf_area = (long long)t.ovf;
  if (stage == 0) {
    // NesterovPlace::init, src/nesterovPlace.cpp:106-119
    s->base_wl_coef = (float)((double)s->init_wl_coef / ((double)s->bin_size_sum * 0.5));
    s->overflow_area = ovf_area;
    s->sum_overflow = overflow_of(s, ovf_area);
    update_wl_coef(s, s->sum_overflow);
    return;
  }
  s->evals++;
  s->wl_grad_sum = t.wl;
  s->dens_grad_sum = t.dens;
  if (bad_sum(t.wl) || bad_sum(t.dens)) {
    if (d.evrec) write_eval_rec(d, s, t, tt, -1, stage - 1, 1, 0.0f, s->sum_overflow);
    s->phase = 2;
    s->code = 2;  // diverged in init -> REPL-2 (src/nesterovPlace.cpp:280-283)
    s->diverged = 1;
    return;
  }
  if (d.evrec)
    write_eval_rec(d, s, t, tt, -1, stage - 1, 0, stage == 2 ? step_of(s, t.dc, t.dg) : 0.0f,
                   stage == 2 ? overflow_of(s, (long long)t.ovf) : s->sum_overflow);
  if (stage == 1) {
    s->prev_hpwl = t.hpwl;  // src/nesterovPlace.cpp:98-99
    // updateInitialPrevSLPCoordi as a step of length initialPrevCoordiUpdateCoef
    s->mode = 0;
    s->dst = s->ring.prev;
    s->step_length = s->initial_prev_coef;
    s->coeff = 0.0f;
    return;
  }
  // stage 2: src/nesterovPlace.cpp:156-170
  s->density_penalty = ((float)t.wl / (float)t.dens) * s->init_density_penalty;
  s->overflow_area = ovf_area;
  s->sum_overflow = overflow_of(s, ovf_area);
  s->step_length = step_of(s, t.dc, t.dg);
  s->cur_a = 1.0f;
  s->min_sum_overflow = 1e30f;
  s->hpwl_with_min_sum_overflow = 1e30f;
  s->is_max_phi_coef_changed = 0;
  s->iter = 0;
  begin_iteration(s);
}

// The controller's scalar logic (thread 0) on the state s, after the totals.
__device__ void control_logic(const RpDev& d, RpState* s, const Totals& t, const TraceTotals& tt) {
  s->evals++;
  s->wl_grad_sum = t.wl;
  s->dens_grad_sum = t.dens;
  s->grad_sum = t.grad;
  s->coord_dist2 = t.dc;
  s->grad_dist2 = t.dg;
  s->overflow_area_d = t.ovf;
  s->hpwl = t.hpwl;
  bool div = bad_sum(t.wl) || bad_sum(t.dens);
  bool accepted = false;
  if (d.evrec)
    write_eval_rec(d, s, t, tt, s->iter, s->bt, div ? 1 : 0, div ? 0.0f : step_of(s, t.dc, t.dg),
                   overflow_of(s, (long long)t.ovf));
  if (!div) {
    float ns = step_of(s, t.dc, t.dg);
    s->last_new_step = ns;
    if ((double)ns > (double)s->step_length * 0.95) {
      s->step_length = ns;
      accepted = true;
    } else {
      s->step_length = ns;
      s->bt++;
    }
  }
  if (!div && !accepted && s->bt < s->max_back_track) return;  // next slot re-evaluates
  // end of the back-tracking loop: src/nesterovPlace.cpp:392-413
  if (!s->is_max_phi_coef_changed && s->sum_overflow < 0.35f) {
    s->is_max_phi_coef_changed = 1;
    s->max_phi_coef = (float)((double)s->max_phi_coef * 0.99);
  }
  if (!div && !accepted) {
    s->code = 3;
    s->phase = 2;
    s->diverged = 1;
    return;
  }
  if (div) {
    s->code = 0;
    s->phase = 2;
    s->diverged = 1;
    return;
  }
  // updateNextIter, src/nesterovPlace.cpp:514-553
  RpRing r = s->ring;
  s->ring.prev = r.cur;
  s->ring.cur = r.next;
  s->ring.next = r.prev;
  s->ring.co_cur = r.co_next;
  s->ring.co_next = r.co_cur;
  long long ovf_area = (long long)t.ovf;
  s->overflow_area = ovf_area;
  s->sum_overflow = overflow_of(s, ovf_area);
  update_wl_coef(s, s->sum_overflow);
  long long hpwl = t.hpwl;
  float scaled = (float)(hpwl - s->prev_hpwl) / s->reference_hpwl;
  float coef = (scaled < 0) ? s->max_phi_coef
                            : (float)((double)s->max_phi_coef * pow((double)s->max_phi_coef, (double)scaled * -1.0));
  coef = (s->min_phi_coef < coef) ? coef : s->min_phi_coef;  // std::max(minPhiCoef, coef)
  s->prev_hpwl = hpwl;
  s->density_penalty *= coef;
  int i = s->iter;
  RpIterRec rec;
  rec.overflow = s->sum_overflow;
  rec.step_length = s->step_length;
  rec.density_penalty = s->density_penalty;
  rec.num_back_track = s->bt + 1;
  rec.hpwl = hpwl;
  d.rec[i] = rec;
  s->iter = i + 1;
  // src/nesterovPlace.cpp:437-462
  if (s->min_sum_overflow > s->sum_overflow) {
    s->min_sum_overflow = s->sum_overflow;
    s->hpwl_with_min_sum_overflow = (float)s->prev_hpwl;
  }
  if (s->sum_overflow < 0.3f && s->sum_overflow - s->min_sum_overflow >= 0.02f &&
      s->hpwl_with_min_sum_overflow * 1.2f < (float)s->prev_hpwl) {
    s->code = 4;
    s->phase = 2;
    s->diverged = 1;
    return;
  }
  if (i > 50 && s->sum_overflow <= s->target_overflow) {
    s->phase = 1;
    return;
  }
  if (s->iter >= s->max_iter) {
    s->phase = 3;
    return;
  }
  begin_iteration(s);
}

// One workgroup: the state comes into LDS in one round (a 4-byte word per
// thread, issued beside the first partial loads), thread 0 runs the logic on
// that copy, and the state goes back in one round, so the scalar logic's
// dependent reads and writes are LDS accesses, not memory round trips.
__device__ void control_main_body(const RpDev& d) {
  __shared__ RpState ls;
  static_assert(sizeof(RpState) % 4 == 0 && sizeof(RpState) / 4 <= RP_CTRL_BLOCK, "state words");
  constexpr int W = sizeof(RpState) / 4;
  if (threadIdx.x < W) reinterpret_cast<int*>(&ls)[threadIdx.x] = reinterpret_cast<const int*>(d.st)[threadIdx.x];
  Totals t;
  reduce_all(d, t);  // its barriers also publish ls
  if (ls.phase != 0) return;  // the same for every thread: the kernels of a stopped run are no-ops
  TraceTotals tt = {0.0, 0.0};
  if (d.evrec) reduce_trace(d, tt);
  if (threadIdx.x == 0) control_logic(d, &ls, t, tt);
  __syncthreads();
  if (threadIdx.x < W) reinterpret_cast<int*>(d.st)[threadIdx.x] = reinterpret_cast<const int*>(&ls)[threadIdx.x];
}


// ---------------------------------------------------------------------------
// kernels: single-instance wrappers (XCD-aware block order) and batched
// wrappers.  A batched launch covers many placement instances: entry b of
// its block table names (instance, logical block) and the instance's device
// pointers come from devs[instance].  Each instance sees exactly the logical
// blocks, block size and reduction order of its single-instance launch, so
// its result is bit-identical alone or in a batch.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(RP_BLOCK) k_step(RpDev d) { step_body(d, xcd_block(blockIdx.x, gridDim.x)); }
// Density scatter in spatial cell order.  The block count is a multiple of
// the XCD count, so xcd_block keeps each XCD on a contiguous range of the
// order.
template <int PER>
__global__ void __launch_bounds__(RP_BLOCK) k_scatter(RpDev d, RpSortBufs sb, int count) {
  sorted_scatter_body<PER>(d, xcd_block(blockIdx.x, gridDim.x), sb, count);
}
static int scatter_per_thread(const RpDev& d) { return d.scatter_per == 2 ? 2 : 1; }
static int scatter_blocks(const RpDev& d) {
  const int cells = RP_BLOCK * scatter_per_thread(d);
  int n = (d.nc + cells - 1) / cells;
  return (n + RP_NUM_XCD - 1) / RP_NUM_XCD * RP_NUM_XCD;
}
static void launch_scatter(const RpDev& d, hipStream_t s, const RpSortBufs& sb, int count) {
  const dim3 grid(scatter_blocks(d));
  if (scatter_per_thread(d) == 2)
    hipLaunchKernelGGL(k_scatter<2>, grid, dim3(RP_BLOCK), 0, s, d, sb, count);
  else
    hipLaunchKernelGGL(k_scatter<1>, grid, dim3(RP_BLOCK), 0, s, d, sb, count);
}
__global__ void __launch_bounds__(RP_BLOCK) k_binfin_rowdct(RpDev d, int from_rho) {
  RP_FFT_DISPATCH(d.nbx, binfin_body, d, blockIdx.x, from_rho)
}
__global__ void __launch_bounds__(RP_BLOCK) k_col_spectral(RpDev d) {
  RP_FFT_DISPATCH(d.nby, col_body, d, xcd_block(blockIdx.x, gridDim.x))
}
__global__ void __launch_bounds__(RP_WIDE) k_col_spectral4(RpDev d) {
  RP_FFT_DISPATCH(d.nby, col4_body, d, xcd_block(blockIdx.x, gridDim.x))
}
__global__ void __launch_bounds__(RP_WIDE) k_row_inverse2(RpDev d, int want_phi) {
  RP_FFT_DISPATCH(d.nbx, rowinv2_body, d, blockIdx.x, want_phi)
}
__global__ void __launch_bounds__(RP_BLOCK) k_row_inverse(RpDev d, int want_phi) {
  RP_FFT_DISPATCH(d.nbx, rowinv_body, d, blockIdx.x, want_phi)
}
__global__ void __launch_bounds__(RP_BLOCK) k_wa(RpDev d, int store_nets) {
  wa_body<RP_WA_CAP>(d, xcd_block(blockIdx.x, gridDim.x), store_nets);
}
// Density scatter and wirelength in one launch (single instance).  Both read
// only the integer centres the step wrote and write disjoint outputs; the
// scatter waits on gathers and memory-side atomics, the wirelength blocks on
// VALU, so sharing the CUs overlaps the two, and one launch gap goes.  The
// LDS is one union of the two tables.  Hardware block b runs on XCD
// b mod 8; in each XCD's sequence of blocks the ns8 scatter blocks are
// spread evenly among the wirelength blocks (a scatter block first), and
// each kind covers the XCD-contiguous logical range xcd_block gives it in a
// launch of its own, so every logical block computes exactly what it does
// there (results are identical).
union ScatterWaLds {
  SortScatterLds s;
  WaLds<RP_WA_CAP> w;
};
template <int PER>
__global__ void __launch_bounds__(RP_BLOCK) __attribute__((amdgpu_waves_per_eu(7)))
k_scatter_wa(RpDev d, RpSortBufs sb, int count, int ns8) {
  __shared__ ScatterWaLds L;
  const int x = blockIdx.x % RP_NUM_XCD, j = blockIdx.x / RP_NUM_XCD;
  const int nwa = max(1, d.nwa), perw = nwa / RP_NUM_XCD, remw = nwa % RP_NUM_XCD;
  const int cnt = ns8 + perw + (x < remw);  // blocks of this XCD
  // scatter blocks among positions [0, j) and [0, j]
  const int c0 = (j * ns8 + cnt - 1) / cnt, c1 = ((j + 1) * ns8 + cnt - 1) / cnt;
  if (c1 > c0)
    sorted_scatter_body<PER>(d, x * ns8 + c0, sb, count, L.s);
  else
    wa_body<RP_WA_CAP>(d, x * perw + min(x, remw) + (j - c1), 0, L.w);
}
__global__ void __launch_bounds__(RP_BLOCK) k_grad(RpDev d) { grad_body(d, xcd_block(blockIdx.x, gridDim.x)); }
__global__ void __launch_bounds__(RP_CTRL_BLOCK) k_control_main(RpDev d) { control_main_body(d); }

template <int PER, int BITS>
__global__ void __launch_bounds__(RP_BLOCK) kb_step_scatter(const RpDev* devs, const int2* tab) {
  int2 e = tab[blockIdx.x];
  if (e.x >= 0) scatter_body<PER, BITS>(devs[e.x], e.y);
}
__global__ void __launch_bounds__(RP_BLOCK) kb_binfin_rowdct(const RpDev* devs, const int2* tab) {
  int2 e = tab[blockIdx.x];
  if (e.x >= 0) RP_FFT_DISPATCH(devs[e.x].nbx, binfin_body, devs[e.x], e.y, 0)
}
__global__ void __launch_bounds__(RP_BLOCK) kb_col_spectral(const RpDev* devs, const int2* tab) {
  int2 e = tab[blockIdx.x];
  if (e.x >= 0) RP_FFT_DISPATCH(devs[e.x].nby, col_body, devs[e.x], e.y)
}
__global__ void __launch_bounds__(RP_BLOCK) kb_row_inverse(const RpDev* devs, const int2* tab) {
  int2 e = tab[blockIdx.x];
  if (e.x >= 0) RP_FFT_DISPATCH(devs[e.x].nbx, rowinv_body, devs[e.x], e.y, 0)
}
template <int CAP>
__global__ void __launch_bounds__(RP_BLOCK) kb_wa(const RpDev* devs, const int2* tab) {
  int2 e = tab[blockIdx.x];
  if (e.x >= 0) wa_body<CAP>(devs[e.x], e.y, 0);
}
__global__ void __launch_bounds__(RP_BLOCK) kb_grad(const RpDev* devs, const int2* tab) {
  int2 e = tab[blockIdx.x];
  if (e.x >= 0) grad_body(devs[e.x], e.y);
}
__global__ void __launch_bounds__(RP_CTRL_BLOCK) kb_control_main(const RpDev* devs, const int2* tab) {
  int2 e = tab[blockIdx.x];
  if (e.x >= 0) control_main_body(devs[e.x]);
}

// ---------------------------------------------------------------------------
// host-callable launch wrappers (C++ linkage, used by rp_host.cpp)
// ---------------------------------------------------------------------------
// FFT-based kernels always run RP_BLOCK threads: the per-thread partial
// sums (overflow) then group identically in single and batched launches.
static int fft_threads(int) { return RP_BLOCK; }

// Column and inverse passes of the single-instance path: the wide forms
// (two transforms side by side per workgroup) from 256 bins per side, where
// they measured faster; the one-pair forms below.
#define RP_WIDE_FROM 256
static void launch_col(const RpDev& d, hipStream_t s) {
  if (d.nbx >= RP_WIDE_FROM && d.nby >= RP_WIDE_FROM)
    hipLaunchKernelGGL(k_col_spectral4, dim3(d.nbx / 4), dim3(RP_WIDE), rp_lds_col4(d.nby), s, d);
  else
    hipLaunchKernelGGL(k_col_spectral, dim3(d.nbx / 2), dim3(fft_threads(d.nby)), rp_lds_col(d.nby), s, d);
}
static void launch_rowinv(const RpDev& d, hipStream_t s, int want_phi) {
  if (d.nbx >= RP_WIDE_FROM && d.nby >= RP_WIDE_FROM)
    hipLaunchKernelGGL(k_row_inverse2, dim3(d.nby / 2), dim3(RP_WIDE), rp_lds_rinv2(d.nbx), s, d, want_phi);
  else
    hipLaunchKernelGGL(k_row_inverse, dim3(d.nby / 2), dim3(fft_threads(d.nbx)), rp_lds_rinv(d.nbx), s, d,
                       want_phi);
}

// One evaluation slot.  ev (optional) holds 2 * RP_NUM_EVAL_KERNELS events:
// ev[2q] and ev[2q + 1] bracket kernel q when it is launched on this path;
// *launched (optional) receives the mask of the kernels that were.
static void launch_scatter_wa(const RpDev& d, hipStream_t s, const RpSortBufs& sb, int count) {
  const int ns = scatter_blocks(d);  // a multiple of the XCD count
  const dim3 grid(ns + std::max(1, d.nwa));
  if (scatter_per_thread(d) == 2)
    hipLaunchKernelGGL(k_scatter_wa<2>, grid, dim3(RP_BLOCK), 0, s, d, sb, count, ns / RP_NUM_XCD);
  else
    hipLaunchKernelGGL(k_scatter_wa<1>, grid, dim3(RP_BLOCK), 0, s, d, sb, count, ns / RP_NUM_XCD);
}

hipError_t rp_launch_eval(const RpDev& d, hipStream_t s, int want_phi, hipEvent_t* ev, const RpMfmaDev* mf,
                          const RpSortBufs* sb, unsigned* launched) {
  const int nb_c = (d.nc + RP_BLOCK - 1) / RP_BLOCK;
  unsigned mask = 0;
  auto beg = [&](int q) {
    if (ev) (void)hipEventRecord(ev[2 * q], s);
    mask |= 1u << q;
  };
  auto end = [&](int q) {
    if (ev) (void)hipEventRecord(ev[2 * q + 1], s);
  };
  beg(RP_K_STEP);
  hipLaunchKernelGGL(k_step, dim3(nb_c), dim3(RP_BLOCK), 0, s, d);
  end(RP_K_STEP);
  beg(RP_K_SCATTER);
  if (d.fuse_scatter_wa)  // the wirelength kernel's work goes into this launch
    launch_scatter_wa(d, s, sb ? *sb : RpSortBufs{}, sb ? 1 : 0);
  else
    launch_scatter(d, s, sb ? *sb : RpSortBufs{}, sb ? 1 : 0);
  end(RP_K_SCATTER);
  beg(RP_K_BINFIN_ROWDCT);
  hipLaunchKernelGGL(k_binfin_rowdct, dim3(d.nby / 2), dim3(fft_threads(d.nbx)), rp_lds_binfin(d.nbx), s, d, 0);
  end(RP_K_BINFIN_ROWDCT);
  if (mf) {
    // matrix-core solve
    beg(RP_K_MFMA_POISSON);
    hipError_t e = rp_launch_poisson_mfma(d, *mf, s, want_phi);
    if (e != hipSuccess) return e;
    end(RP_K_MFMA_POISSON);
  } else {
    beg(RP_K_COL_SPECTRAL);
    launch_col(d, s);
    end(RP_K_COL_SPECTRAL);
    beg(RP_K_ROW_INVERSE);
    launch_rowinv(d, s, want_phi);
    end(RP_K_ROW_INVERSE);
  }
  if (!d.fuse_scatter_wa) {
    beg(RP_K_WA);
    hipLaunchKernelGGL(k_wa, dim3(std::max(1, d.nwa)), dim3(RP_BLOCK), 0, s, d, 0);
    end(RP_K_WA);
  }
  beg(RP_K_GRAD);
  hipLaunchKernelGGL(k_grad, dim3(nb_c), dim3(RP_BLOCK), 0, s, d);
  end(RP_K_GRAD);
  beg(RP_K_CONTROL);
  hipLaunchKernelGGL(k_control_main, dim3(1), dim3(RP_CTRL_BLOCK), 0, s, d);
  end(RP_K_CONTROL);
  if (launched) *launched = mask;
  return hipGetLastError();
}

hipError_t rp_launch_ctrl_init(const RpDev& d, hipStream_t s, int stage) {
  hipLaunchKernelGGL(k_control_init, dim3(1), dim3(RP_CTRL_BLOCK), 0, s, d, stage);
  return hipGetLastError();
}

// Split stages for the stage-level test entry points.
hipError_t rp_launch_scatter(const RpDev& d, hipStream_t s) {
  hipLaunchKernelGGL(k_step, dim3((d.nc + RP_BLOCK - 1) / RP_BLOCK), dim3(RP_BLOCK), 0, s, d);
  launch_scatter(d, s, RpSortBufs{}, 0);
  return hipGetLastError();
}

hipError_t rp_launch_poisson(const RpDev& d, hipStream_t s, int want_phi, int from_rho, const RpMfmaDev* mf) {
  size_t shx = rp_lds_binfin(d.nbx);
  hipLaunchKernelGGL(k_binfin_rowdct, dim3(d.nby / 2), dim3(fft_threads(d.nbx)), shx, s, d, from_rho);
  if (mf) return rp_launch_poisson_mfma(d, *mf, s, want_phi);
  launch_col(d, s);
  launch_rowinv(d, s, want_phi);
  return hipGetLastError();
}

hipError_t rp_launch_wa_grad(const RpDev& d, hipStream_t s, int store_nets) {
  hipLaunchKernelGGL(k_wa, dim3(std::max(1, d.nwa)), dim3(RP_BLOCK), 0, s, d, store_nets);
  hipLaunchKernelGGL(k_grad, dim3((d.nc + RP_BLOCK - 1) / RP_BLOCK), dim3(RP_BLOCK), 0, s, d);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// spatial cell order for the density scatter (counting sort by tile)
// ---------------------------------------------------------------------------
// Exclusive scan of the RP_SORT_TILES tile counts into offsets (one
// workgroup of RP_CTRL_BLOCK threads, each owning a contiguous run of
// tiles); the counts are cleared for the next refresh.
__global__ void __launch_bounds__(RP_CTRL_BLOCK) k_sort_scan(RpSortBufs sb) {
  constexpr int PER = RP_SORT_TILES / RP_CTRL_BLOCK;
  __shared__ int sh[RP_CTRL_BLOCK];
  int base = threadIdx.x * PER;
  int v[PER];
  int run = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    v[k] = sb.cnt[base + k];
    sb.cnt[base + k] = 0;
    run += v[k];
  }
  sh[threadIdx.x] = run;
  __syncthreads();
  for (int o = 1; o < RP_CTRL_BLOCK; o <<= 1) {  // Hillis-Steele inclusive scan of the run totals
    int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int acc = sh[threadIdx.x] - run;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    sb.off[base + k] = acc;
    acc += v[k];
  }
}

// Second pass: each cell goes to its tile's offset + its rank.  The counts
// came from the same integer centres (no step ran in between), so this
// writes a permutation of 0 .. Nc-1.
__global__ void __launch_bounds__(RP_BLOCK) k_sort_place(RpDev d, RpSortBufs sb) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.nc) return;
  sb.sord[sb.off[tile_key(d, d.ic[c], sb.tile_shift)] + sb.rank[c]] = c;
}

int rp_sort_tile_shift(int nbx, int nby) {
  int m = std::max(nbx, nby), lg = 0;
  while ((1 << lg) < m) lg++;
  return std::max(0, lg - 7);
}

hipError_t rp_launch_sort_refresh(const RpDev& d, const RpSortBufs& b, hipStream_t s) {
  if (d.nc <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(RP_CTRL_BLOCK), 0, s, b);
  hipLaunchKernelGGL(k_sort_place, dim3((d.nc + RP_BLOCK - 1) / RP_BLOCK), dim3(RP_BLOCK), 0, s, d, b);
  return hipGetLastError();
}

__global__ void k_gather_phase(const RpDev* devs, int n, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = devs[i].st->phase;
}

hipError_t rp_launch_gather_phase(const RpDev* devs, int n, int* out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_phase, dim3((n + 255) / 256), dim3(256), 0, s, devs, n, out);
  return hipGetLastError();
}

// One evaluation step of every instance in the tables (same kernels, block
// sizes and per-instance logical blocks as rp_launch_eval).
// Tile-shape variants of the batched scatter and wirelength kernels.
hipError_t rp_launch_batch_variant(int kernel, int variant, const RpDev* devs, const int2* tab, int nblk,
                                   hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  if (kernel == RP_K_SCATTER) {
    if (variant == 0)
      hipLaunchKernelGGL((kb_step_scatter<1, 10>), dim3(nblk), dim3(RP_BLOCK), 0, s, devs, tab);
    else if (variant == 1)
      hipLaunchKernelGGL((kb_step_scatter<2, 11>), dim3(nblk), dim3(RP_BLOCK), 0, s, devs, tab);
    else
      hipLaunchKernelGGL((kb_step_scatter<4, 12>), dim3(nblk), dim3(RP_BLOCK), 0, s, devs, tab);
  } else if (kernel == RP_K_WA) {
    if (variant == 0)
      hipLaunchKernelGGL((kb_wa<256>), dim3(nblk), dim3(RP_BLOCK), 0, s, devs, tab);
    else if (variant == 1)
      hipLaunchKernelGGL((kb_wa<512>), dim3(nblk), dim3(RP_BLOCK), 0, s, devs, tab);
    else
      hipLaunchKernelGGL((kb_wa<1024>), dim3(nblk), dim3(RP_BLOCK), 0, s, devs, tab);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t rp_launch_eval_batch(const RpBatchLaunch& b, hipStream_t s) {
  const int* n = b.nblk;
  for (int v = 0; v < RP_SC_VARIANTS; v++)
    if (b.nblk_sc[v]) (void)rp_launch_batch_variant(RP_K_SCATTER, v, b.devs, b.tab_sc[v], b.nblk_sc[v], s);
  if (n[RP_K_BINFIN_ROWDCT])
    hipLaunchKernelGGL(kb_binfin_rowdct, dim3(n[RP_K_BINFIN_ROWDCT]), dim3(RP_BLOCK), b.lds_binfin, s, b.devs,
                       b.tab[RP_K_BINFIN_ROWDCT]);
  if (n[RP_K_COL_SPECTRAL])
    hipLaunchKernelGGL(kb_col_spectral, dim3(n[RP_K_COL_SPECTRAL]), dim3(RP_BLOCK), b.lds_col, s, b.devs,
                       b.tab[RP_K_COL_SPECTRAL]);
  if (n[RP_K_ROW_INVERSE])
    hipLaunchKernelGGL(kb_row_inverse, dim3(n[RP_K_ROW_INVERSE]), dim3(RP_BLOCK), b.lds_rinv, s, b.devs,
                       b.tab[RP_K_ROW_INVERSE]);
  for (int v = 0; v < RP_WA_VARIANTS; v++)
    if (b.nblk_wa[v]) (void)rp_launch_batch_variant(RP_K_WA, v, b.devs, b.tab_wa[v], b.nblk_wa[v], s);
  if (n[RP_K_GRAD]) hipLaunchKernelGGL(kb_grad, dim3(n[RP_K_GRAD]), dim3(RP_BLOCK), 0, s, b.devs, b.tab[RP_K_GRAD]);
  if (n[RP_K_CONTROL])
    hipLaunchKernelGGL(kb_control_main, dim3(n[RP_K_CONTROL]), dim3(RP_CTRL_BLOCK), 0, s, b.devs,
                       b.tab[RP_K_CONTROL]);
  return hipGetLastError();
}

// Loads this file's code object on the current device (see rp_ip_load_code).
hipError_t rp_kernels_load_code() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_grad));
}
