// rp_dct_mfma.hip -- the Poisson solve (FFT::doFFT, src/fft.cpp:112-177) as
// dense transform-matrix GEMMs on CDNA4 matrix cores, split-BF16 operands
// (value = hi + lo, both bf16) and FP32 accumulation.
//
// Transform (x = column index i < nx, y = row index j < ny, rho[j][i]):
//   A[v][u]  = w(u,v) sum_j sum_i cos(pi v (2j+1)/2ny) rho[j][i] cos(pi u (2i+1)/2nx)
//   w(u,v)   = 4/(nx ny), halved on u = 0 and on v = 0
//   P[v][u]  = A[v][u] / (wx_u^2 + wy_v^2), P[0][0] = 0,
//              wx_u = pi u / nx, wy_v = pi v / ny * (binSizeY / binSizeX)
//   phi[j][i] = sum_v sum_u P[v][u] cos(pi v (2j+1)/2ny) cos(pi u (2i+1)/2nx)
//   Ex[j][i]  = sum_v sum_u P[v][u] wx_u cos(pi v (2j+1)/2ny) sin(pi u (2i+1)/2nx)
//   Ey[j][i]  = sum_v sum_u P[v][u] wy_v sin(pi v (2j+1)/2ny) cos(pi u (2i+1)/2nx)
// as five "NT" products C[m][n] = sum_k A[m][k] B[n][k] whose A is always a
// constant transform matrix and whose B is the data, so both operands are
// read along k and no transpose pass is needed:
//   G1  T1t[u][j]          = Cx[u][i]            . rho[j][i]
//   G2  P[v][u]            = Cy[v][j]            . T1t[u][j]     (spectral scaling after it)
//   G3  [R1t; R2t][i][v]   = [Cxt; wx Sxt][i][u] . P[v][u]
//   G4  [phi | Ex][j][.]   = Cyt[j][v]           . [R1t; R2t][.][v]
//   G5  Ey[j][i]           = (wy Syt)[j][v]      . R1t[i][v]
//
// Split products: with A = Ahi + Alo and B = Bhi + Blo (bf16 each),
// A.B ~ Ahi.Bhi + Ahi.Blo + Alo.Bhi (the lo.lo term is below the FP32
// rounding of the accumulation).  That is ONE plain bf16 GEMM with FP32
// accumulation over concatenated operands, A' = [Ahi | Ahi | Alo] and
// B' = [Bhi | Blo | Bhi] (K' = 3K), run by this file's k_gemm_nt on the
// matrix cores.  The constant A' are built once per set-up; the data
// operands B' are written by this file's kernels between the GEMMs (split,
// spectral scaling, the scatter of the last products into phi and the field),
// which also add the split-K partials of the product before them.  Per
// evaluation at nx = ny = N the products are N^3 + N^3 + 2 N^3 + N^3 + N^3 =
// 6 N^3 multiply-adds (7 N^3 with phi), each over the 3 split components.
// RP_MFMA_GEMM=hipblaslt runs the same products through hipBLASLt instead
// (the set-up times every algorithm its heuristic offers and keeps the
// fastest), as a reference point for k_gemm_nt.
#include <hipblaslt/hipblaslt.h>

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "rp_ctx.h"

typedef __bf16 rp_bf16;

namespace {

constexpr int MEAN_PARTS = 1024;  // blocks of the mean's first pass
constexpr size_t kWorkspace = (size_t)64 << 20;
constexpr int kAlgos = 8;  // heuristic candidates timed per GEMM

enum { G1 = 0, G2, G3, G4_PHI, G4, G5, NGEMM };

struct Gemm {
  hipblasLtMatmulDesc_t desc = nullptr;
  hipblasLtMatrixLayout_t lb = nullptr, la = nullptr, lc = nullptr;
  hipblasLtMatmulAlgo_t algo{};
  int M = 0, N = 0, K = 0;  // C[M][N] = sum over 3K of A'[M][3K] B'[N][3K]
  float us = 0;             // measured time of the chosen algorithm (hipBLASLt)
  int tile = 0;             // k_gemm_nt tile: 256 (256 x 256, 8 waves), 128 or 64 (4 waves)
  int splitk = 1;           // k slices of k_gemm_nt (partial products summed by the consumer)
};

struct Plan {
  bool lt = false;  // products through hipBLASLt (RP_MFMA_GEMM=hipblaslt) instead of k_gemm_nt
  hipblasLtHandle_t h = nullptr;
  Gemm g[NGEMM];
  DevBuf<uint16_t> a[5];                  // A' of G1, G2, G3, G4, G5 (row-major M x 3K)
  DevBuf<uint16_t> rho3, t13, p3, r3;     // B' of G1, G2, G3, G4/G5 (row-major N x 3K)
  DevBuf<float> c;                        // FP32 product (largest: 2N x N)
  DevBuf<float> wx2, wy2;
  DevBuf<double> sums;                    // mean of rho: partials, then the mean
  DevBuf<char> ws;
  ~Plan() {
    for (Gemm& x : g) {
      if (x.desc) hipblasLtMatmulDescDestroy(x.desc);
      if (x.la) hipblasLtMatrixLayoutDestroy(x.la);
      if (x.lb) hipblasLtMatrixLayoutDestroy(x.lb);
      if (x.lc) hipblasLtMatrixLayoutDestroy(x.lc);
    }
    if (h) hipblasLtDestroy(h);
  }
};

__device__ __forceinline__ void split_bf16(float v, rp_bf16& h, rp_bf16& l) {
  h = (rp_bf16)v;
  l = (rp_bf16)(v - (float)h);
}

// Row n of a B' operand: [hi | lo | hi] of n values x[0..k).
__device__ __forceinline__ void put_split3(rp_bf16* row, int k, int i, float v) {
  rp_bf16 h, l;
  split_bf16(v, h, l);
  row[i] = h;
  row[k + i] = l;
  row[2 * k + i] = h;
}

// ---------------------------------------------------------------------------
// Split-BF16 products on the matrix cores: C[m][n] = sum_k A'[m][k] B'[n][k]
// (both operands row-major along k), FP32 accumulation, v_mfma_f32_32x32x16_bf16.
//
// A workgroup of NW waves computes a BM x BN tile; the waves form a
// 2 x (NW / 2) grid of (BM / 2) x (BN / (NW / 2)) sub-tiles of 32 x 32 MFMA
// tiles.  K is walked in BK = 64 slices through two LDS stages filled by
// LDS-DMA loads (global_load_lds_dwordx4: each lane's 16 bytes land at
// M0 + 16 lane, so the lanes fetch the global chunks in the order of the
// swizzled LDS layout): the next slice streams in while the current one is
// multiplied.  A row of a slice is 8 chunks of 16 bytes, stored at chunk
// kc ^ (row % 8), so the 32 rows a fragment load touches spread over all
// banks.  The fragment of an MFMA operand is 8 consecutive k of one row
// (lane l: row l % 32, k-half l / 32), read with one ds_read_b128.
// With split-K the launch's blocks cover (k slice, tile) pairs, k slice
// outermost, and write their partial products to C + ks * M * N; the
// consumer adds the partials in ks order.  A tile's blocks follow an
// XCD-aware order (hardware block b runs on XCD b % 8): XCD x takes a
// (tiles_m / 4) x (tiles_n / 2) sub-grid, so its blocks share A and B slices
// in its L2.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GBK = 64;  // k slice

__device__ __forceinline__ int gswz(int row, int kc) { return row * 8 + (kc ^ (row & 7)); }

// vmcnt <= 0 (expcnt, lgkmcnt not waited)
__device__ __forceinline__ void wait_vm0() { __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8)); }

template <int BM, int BN, int NW>
__global__ void __launch_bounds__(64 * NW) k_gemm_nt(const rp_bf16* __restrict__ A, const rp_bf16* __restrict__ B,
                                                     float* __restrict__ C, int M, int N, int K, int splitk) {
  constexpr int WM = BM / 2, WN = BN / (NW / 2), TM = WM / 32, TN = WN / 32;
  constexpr int QA = BM * 8 / 64 / NW, QB = BN * 8 / 64 / NW;  // LDS-DMA loads per wave and slice
  static_assert(TM >= 1 && TN >= 1 && QA >= 1 && QB >= 1, "tile too small for the wave grid");
  __shared__ uint4 lds[2][(BM + BN) * 8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tm = M / BM, tn = N / BN, T = tm * tn;
  const int ks = blockIdx.x / T, t = blockIdx.x % T;
  int bm, bn;
  if (tm % 4 == 0 && tn % 2 == 0 && T % 8 == 0) {
    const int x = t % 8, j = t / 8, sm = tm / 4, sn = tn / 2;
    bm = (x / 2) * sm + j / sn;
    bn = (x % 2) * sn + j % sn;
  } else {
    bm = t / tn;
    bn = t % tn;
  }
  const int KT = K / GBK / splitk;
  const rp_bf16* Ab = A + (size_t)bm * BM * K + (size_t)ks * KT * GBK;
  const rp_bf16* Bb = B + (size_t)bn * BN * K + (size_t)ks * KT * GBK;
  f32x16 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;
  const int wm = (wave / (NW / 2)) * WM, wn = (wave % (NW / 2)) * WN;
  // LDS chunk p of a stage (A rows, then B rows) holds global chunk
  // (row p / 8, kc) with gswz(row, kc) = p
  auto issue = [&](int kt, int stage) {
#pragma unroll
    for (int i = 0; i < QA; i++) {
      const int q = i * NW + wave, p = q * 64 + lane, row = p >> 3, kc = (p & 7) ^ (row & 7);
      __builtin_amdgcn_global_load_lds(
          (const __attribute__((address_space(1))) void*)(Ab + (size_t)row * K + kt * GBK + kc * 8),
          (__attribute__((address_space(3))) void*)&lds[stage][q * 64], 16, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < QB; i++) {
      const int q = i * NW + wave, p = q * 64 + lane, row = p >> 3, kc = (p & 7) ^ (row & 7);
      __builtin_amdgcn_global_load_lds(
          (const __attribute__((address_space(1))) void*)(Bb + (size_t)row * K + kt * GBK + kc * 8),
          (__attribute__((address_space(3))) void*)&lds[stage][BM * 8 + q * 64], 16, 0, 0);
    }
  };
#define GFRAG(FA, FB, buf, s)                                                                  \
  {                                                                                            \
    const int kc = 2 * (s) + (lane >> 5);                                                      \
    _Pragma("unroll") for (int i = 0; i < TM; i++)                                             \
      FA[i] = __builtin_bit_cast(bf16x8, lds[buf][gswz(wm + 32 * i + (lane & 31), kc)]);      \
    _Pragma("unroll") for (int j = 0; j < TN; j++)                                             \
      FB[j] = __builtin_bit_cast(bf16x8, lds[buf][gswz(BM + wn + 32 * j + (lane & 31), kc)]); \
  }
#define GMF(FA, FB)                                                                            \
  _Pragma("unroll") for (int i = 0; i < TM; i++)                                               \
    _Pragma("unroll") for (int j = 0; j < TN; j++)                                             \
      acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(FA[i], FB[j], acc[i][j], 0, 0, 0);
  if (KT > 0) issue(0, 0);
  for (int kt = 0; kt < KT; kt++) {
    const int buf = kt & 1;
    wait_vm0();  // this wave's loads of slice kt have landed in LDS
    // a barrier without the release fence of __syncthreads (that fence would
    // also wait for loads still in flight); the empty asm keeps the LDS
    // accesses on their side of it
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (kt + 1 < KT) issue(kt + 1, buf ^ 1);  // the other stage was read by slice kt - 1, before the barrier
    bf16x8 xa[TM], xb[TN], ya[TM], yb[TN];
    GFRAG(xa, xb, buf, 0)
    GFRAG(ya, yb, buf, 1)
    GMF(xa, xb)
    GFRAG(xa, xb, buf, 2)
    GMF(ya, yb)
    GFRAG(ya, yb, buf, 3)
    GMF(xa, xb)
    GMF(ya, yb)
  }
#undef GFRAG
#undef GMF
  float* Cp = C + (size_t)ks * M * N;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int m0 = bm * BM + wm + 32 * i, n0 = bn * BN + wn + 32 * j;
#pragma unroll
      for (int r = 0; r < 16; r++)
        Cp[(size_t)(m0 + (r / 4) * 8 + (lane / 32) * 4 + r % 4) * N + n0 + (lane & 31)] = acc[i][j][r];
    }
}

// sum of the split-K partials of element k, in partial order
__device__ __forceinline__ float psum_c(const float* __restrict__ c, size_t k, int nsplit, size_t stride) {
  float v = c[k];
  for (int s = 1; s < nsplit; s++) v += c[k + (size_t)s * stride];
  return v;
}

// Mean of rho (fixed-order FP64 partials, then one block).  Subtracting a
// constant from rho changes only the u = v = 0 coefficient, which the
// spectral step discards, so the transform of (rho - mean) is exact; it keeps
// the 16-bit rounding of the large mean out of the low modes that the
// spectral division amplifies.
__device__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
  return r;
}

__global__ void __launch_bounds__(256) k_rho_parts(const float* __restrict__ x, size_t n, double* parts) {
  __shared__ double sh[4];
  double s = 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    s += (double)x[k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_rho_mean(const double* parts, size_t n, double* mean) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int k = threadIdx.x; k < MEAN_PARTS; k += blockDim.x) s += parts[k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) mean[0] = s / (double)n;
}

// rho (FP32, [ny][nx]) - mean -> B' of G1 (rows j, k = i)
__global__ void __launch_bounds__(256) k_split_rho(const float* __restrict__ x, const double* __restrict__ mean,
                                                   rp_bf16* __restrict__ out, int nx, int ny) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)nx * ny) return;
  int j = (int)(k / nx), i = (int)(k % nx);
  put_split3(out + (size_t)j * 3 * nx, nx, i, (float)((double)x[k] - mean[0]));
}

// FP32 product C[m][n] -> B' rows m (k = n); with `spectral` the scaling of
// FFT::doFFT first: w(u,v) / (wx_u^2 + wy_v^2), m = v, n = u, 0 at u = v = 0.
__global__ void __launch_bounds__(256) k_split_c(const float* __restrict__ c, rp_bf16* __restrict__ out, int M,
                                                 int N, int spectral, const float* __restrict__ wx2,
                                                 const float* __restrict__ wy2, float wscale, int nsplit) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)M * N) return;
  int m = (int)(k / N), n = (int)(k % N);
  float v = psum_c(c, k, nsplit, (size_t)M * N);
  if (spectral) {
    float s = 0.0f;
    if (m != 0 || n != 0) {
      float w = wscale;
      if (n == 0) w *= 0.5f;
      if (m == 0) w *= 0.5f;
      s = w / (wx2[n] + wy2[m]);
    }
    v = v * s;
  }
  put_split3(out + (size_t)m * 3 * N, N, n, v);
}

// G4's [phi | Ex][j][c] (columns from col0) and G5's Ey[j][i] into phi and the field
__global__ void __launch_bounds__(256) k_phi_ex(const float* __restrict__ c, float* __restrict__ phi,
                                                float2* __restrict__ field, int nx, int ny, int ncol, int col0,
                                                int nsplit) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)ny * ncol) return;
  int j = (int)(k / ncol), cc = col0 + (int)(k % ncol);
  const float v = psum_c(c, k, nsplit, (size_t)ny * ncol);
  if (cc < nx)
    phi[(size_t)j * nx + cc] = v;
  else
    field[(size_t)j * nx + (cc - nx)].x = v;
}

__global__ void __launch_bounds__(256) k_ey(const float* __restrict__ c, float2* __restrict__ field, size_t n,
                                            int nsplit) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) field[k].y = psum_c(c, k, nsplit, n);
}

// bf16 round-to-nearest-even of a float (host)
uint16_t bf16_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_float(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// A' = [hi | hi | lo] of a row-major M x K matrix
std::vector<uint16_t> split3_const(const std::vector<double>& v, int M, int K) {
  std::vector<uint16_t> out((size_t)M * 3 * K);
  for (int m = 0; m < M; m++)
    for (int k = 0; k < K; k++) {
      float f = (float)v[(size_t)m * K + k];
      uint16_t h = bf16_bits(f), l = bf16_bits(f - bf16_float(h));
      uint16_t* row = &out[(size_t)m * 3 * K];
      row[k] = h;
      row[K + k] = h;
      row[2 * K + k] = l;
    }
  return out;
}

// cos / sin of pi * num / den with the angle reduced exactly in integers
double cospi_frac(long long num, long long den) {
  long long r = num % (2 * den);
  return std::cos(M_PI * (double)r / (double)den);
}
double sinpi_frac(long long num, long long den) {
  long long r = num % (2 * den);
  return std::sin(M_PI * (double)r / (double)den);
}

#define BLT(c, x)                                                                        \
  do {                                                                                   \
    hipblasStatus_t e_ = (x);                                                            \
    if (e_ != HIPBLAS_STATUS_SUCCESS) return fail(c, RP_ERR_HIP, "%s: status %d", #x, (int)e_); \
  } while (0)

// Row-major C[M][N] = A'[M][3K] . B'[N][3K]^T in hipBLASLt's column-major
// terms: D (N x M, ld N) = op(X) op(Y) with X = B' seen as 3K x N (transposed)
// and Y = A' seen as 3K x M.
int make_gemm(rp_ctx* c, Plan& P, Gemm& g, int M, int N, int K) {
  g.M = M, g.N = N, g.K = K;
  const int K3 = 3 * K;
  BLT(c, hipblasLtMatmulDescCreate(&g.desc, HIPBLAS_COMPUTE_32F, HIP_R_32F));
  hipblasOperation_t ta = HIPBLAS_OP_T, tb = HIPBLAS_OP_N;
  BLT(c, hipblasLtMatmulDescSetAttribute(g.desc, HIPBLASLT_MATMUL_DESC_TRANSA, &ta, sizeof ta));
  BLT(c, hipblasLtMatmulDescSetAttribute(g.desc, HIPBLASLT_MATMUL_DESC_TRANSB, &tb, sizeof tb));
  BLT(c, hipblasLtMatrixLayoutCreate(&g.lb, HIP_R_16BF, K3, N, K3));
  BLT(c, hipblasLtMatrixLayoutCreate(&g.la, HIP_R_16BF, K3, M, K3));
  BLT(c, hipblasLtMatrixLayoutCreate(&g.lc, HIP_R_32F, N, M, N));
  hipblasLtMatmulPreference_t pref;
  BLT(c, hipblasLtMatmulPreferenceCreate(&pref));
  size_t wsz = kWorkspace;
  hipblasStatus_t st =
      hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsz, sizeof wsz);
  hipblasLtMatmulHeuristicResult_t res[kAlgos];
  int n = 0;
  if (st == HIPBLAS_STATUS_SUCCESS)
    st = hipblasLtMatmulAlgoGetHeuristic(P.h, g.desc, g.lb, g.la, g.lc, g.lc, pref, kAlgos, res, &n);
  hipblasLtMatmulPreferenceDestroy(pref);
  BLT(c, st);
  if (n <= 0) return fail(c, RP_ERR_HIP, "hipBLASLt has no bf16 GEMM for %d x %d x %d", M, N, K3);
  g.algo = res[0].algo;
  return 0;
}

hipblasStatus_t run_gemm(const Plan& P, const Gemm& g, const uint16_t* a, const uint16_t* b, float* out,
                         hipStream_t s, const hipblasLtMatmulAlgo_t* algo = nullptr) {
  const float alpha = 1.0f, beta = 0.0f;
  return hipblasLtMatmul(P.h, g.desc, &alpha, b, g.lb, a, g.la, &beta, out, g.lc, out, g.lc,
                         algo ? algo : &g.algo, P.ws.p, kWorkspace, s);
}

// Times every heuristic candidate of g on its real operands, keeps the fastest.
int tune_gemm(rp_ctx* c, Plan& P, Gemm& g, const uint16_t* a, const uint16_t* b, hipStream_t s) {
  hipblasLtMatmulPreference_t pref;
  BLT(c, hipblasLtMatmulPreferenceCreate(&pref));
  size_t wsz = kWorkspace;
  hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsz, sizeof wsz);
  hipblasLtMatmulHeuristicResult_t res[kAlgos];
  int n = 0;
  hipblasStatus_t st = hipblasLtMatmulAlgoGetHeuristic(P.h, g.desc, g.lb, g.la, g.lc, g.lc, pref, kAlgos, res, &n);
  hipblasLtMatmulPreferenceDestroy(pref);
  BLT(c, st);
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  float best = 1e30f;
  for (int k = 0; k < n; k++) {
    if (run_gemm(P, g, a, b, P.c.p, s, &res[k].algo) != HIPBLAS_STATUS_SUCCESS) continue;  // warm
    HIPCHK(c, hipEventRecord(e0, s));
    for (int r = 0; r < 3; r++) BLT(c, run_gemm(P, g, a, b, P.c.p, s, &res[k].algo));
    HIPCHK(c, hipEventRecord(e1, s));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    if (ms / 3 < best) {
      best = ms / 3;
      g.algo = res[k].algo;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g.us = best * 1e3f;
  return 0;
}

inline dim3 grid_of(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// k_gemm_nt shape of a product: the largest tile that divides M and N, and
// k slices (1, 2 or 4, dividing K' / 64, at most cmax / (M N) of them) until
// the launch has at least 256 blocks (one per CU)
int plan_hand(rp_ctx* c, Gemm& g, size_t cmax) {
  const int K3 = 3 * g.K;
  g.tile = g.M % 256 == 0 && g.N % 256 == 0 ? 256 : g.M % 128 == 0 && g.N % 128 == 0 ? 128
           : g.M % 64 == 0 && g.N % 64 == 0 ? 64 : 0;
  if (!g.tile || K3 % GBK)
    return fail(c, RP_ERR_ARG, "matrix-core transform: bin grids of at least 64 bins a side (multiples of 64)");
  const long tiles = (long)(g.M / g.tile) * (g.N / g.tile);
  const int kt = K3 / GBK;
  g.splitk = 1;
  while (tiles * g.splitk < 256 && g.splitk < 4 && kt % (2 * g.splitk) == 0 &&
         (size_t)2 * g.splitk * g.M * g.N <= cmax)
    g.splitk *= 2;
  return 0;
}

hipError_t run_hand(const Gemm& g, const uint16_t* a, const uint16_t* b, float* out, hipStream_t s) {
  const int K3 = 3 * g.K;
  const unsigned blocks = (unsigned)((g.M / g.tile) * (g.N / g.tile) * g.splitk);
  const rp_bf16* A = reinterpret_cast<const rp_bf16*>(a);
  const rp_bf16* B = reinterpret_cast<const rp_bf16*>(b);
  if (g.tile == 256)
    hipLaunchKernelGGL((k_gemm_nt<256, 256, 8>), dim3(blocks), dim3(512), 0, s, A, B, out, g.M, g.N, K3, g.splitk);
  else if (g.tile == 128)
    hipLaunchKernelGGL((k_gemm_nt<128, 128, 4>), dim3(blocks), dim3(256), 0, s, A, B, out, g.M, g.N, K3, g.splitk);
  else
    hipLaunchKernelGGL((k_gemm_nt<64, 64, 4>), dim3(blocks), dim3(256), 0, s, A, B, out, g.M, g.N, K3, g.splitk);
  return hipGetLastError();
}

// one product, through k_gemm_nt or hipBLASLt; nsplit: partials in out
hipError_t product(const Plan& P, const Gemm& g, const uint16_t* a, const uint16_t* b, float* out, hipStream_t s,
                   int* nsplit) {
  if (!P.lt) {
    *nsplit = g.splitk;
    return run_hand(g, a, b, out, s);
  }
  *nsplit = 1;
  return run_gemm(P, g, a, b, out, s) == HIPBLAS_STATUS_SUCCESS ? hipSuccess : hipErrorUnknown;
}

}  // namespace

RpMfma::~RpMfma() { delete static_cast<Plan*>(plan); }

// GEMM plans and constant split-BF16 transform matrices (once per set-up).
int rp_mfma_setup(rp_ctx* c) {
  const int nx = c->nbx, ny = c->nby;
  hipStream_t st = c->stream;
  RpMfma& m = c->mfma;
  delete static_cast<Plan*>(m.plan);
  m.plan = nullptr;
  m.ready = false;
  Plan* P = new Plan();
  m.plan = P;
  const char* impl = getenv("RP_MFMA_GEMM");
  P->lt = impl && strcmp(impl, "hipblaslt") == 0;
  if (P->lt) BLT(c, hipblasLtCreate(&P->h));
  const double asp = (double)c->bsy / (double)c->bsx;
  auto put = [&](const std::vector<double>& v, int M, int K, DevBuf<uint16_t>& dst) -> int {
    HIPCHK(c, dupload(dst, split3_const(v, M, K), st));
    HIPCHK(c, hipStreamSynchronize(st));
    return 0;
  };
  std::vector<double> M;
  // G1: Cx[u][i]
  M.assign((size_t)nx * nx, 0.0);
  for (int u = 0; u < nx; u++)
    for (int i = 0; i < nx; i++) M[(size_t)u * nx + i] = cospi_frac((long long)u * (2 * i + 1), 2LL * nx);
  if (int r = put(M, nx, nx, P->a[0])) return r;
  // G2: Cy[v][j]
  M.assign((size_t)ny * ny, 0.0);
  for (int v = 0; v < ny; v++)
    for (int j = 0; j < ny; j++) M[(size_t)v * ny + j] = cospi_frac((long long)v * (2 * j + 1), 2LL * ny);
  if (int r = put(M, ny, ny, P->a[1])) return r;
  // G3: [Cxt; wx Sxt][i][u]
  M.assign((size_t)2 * nx * nx, 0.0);
  for (int i = 0; i < nx; i++)
    for (int u = 0; u < nx; u++) {
      M[(size_t)i * nx + u] = cospi_frac((long long)u * (2 * i + 1), 2LL * nx);
      M[(size_t)(nx + i) * nx + u] = (M_PI * u / nx) * sinpi_frac((long long)u * (2 * i + 1), 2LL * nx);
    }
  if (int r = put(M, 2 * nx, nx, P->a[2])) return r;
  // G4: Cyt[j][v]
  M.assign((size_t)ny * ny, 0.0);
  for (int j = 0; j < ny; j++)
    for (int v = 0; v < ny; v++) M[(size_t)j * ny + v] = cospi_frac((long long)v * (2 * j + 1), 2LL * ny);
  if (int r = put(M, ny, ny, P->a[3])) return r;
  // G5: (wy Syt)[j][v]
  for (int j = 0; j < ny; j++)
    for (int v = 0; v < ny; v++)
      M[(size_t)j * ny + v] = (M_PI * v / ny * asp) * sinpi_frac((long long)v * (2 * j + 1), 2LL * ny);
  if (int r = put(M, ny, ny, P->a[4])) return r;
  std::vector<float> wx2(nx), wy2(ny);
  for (int u = 0; u < nx; u++) wx2[u] = (float)((M_PI * u / nx) * (M_PI * u / nx));
  for (int v = 0; v < ny; v++) wy2[v] = (float)((M_PI * v / ny * asp) * (M_PI * v / ny * asp));
  HIPCHK(c, dupload(P->wx2, wx2, st));
  HIPCHK(c, dupload(P->wy2, wy2, st));
  const size_t B = (size_t)nx * ny;
  HIPCHK(c, dalloc(P->rho3, 3 * B));
  HIPCHK(c, dalloc(P->t13, 3 * B));
  HIPCHK(c, dalloc(P->p3, 3 * B));
  HIPCHK(c, dalloc(P->r3, 6 * B));
  // the FP32 products: 2 B values (G3 / G4 with phi), or the split-K partials
  const size_t cmax = 4 * B;
  HIPCHK(c, dalloc(P->c, cmax));
  HIPCHK(c, dalloc(P->sums, MEAN_PARTS + 1));
  for (auto* buf : {&P->rho3, &P->t13, &P->p3, &P->r3})
    HIPCHK(c, hipMemsetAsync(buf->p, 0, buf->n * sizeof(uint16_t), st));
  const int shapes[NGEMM][3] = {{nx, ny, nx}, {ny, nx, ny}, {2 * nx, ny, nx}, {ny, 2 * nx, ny}, {ny, nx, ny},
                                {ny, nx, ny}};
  for (int k = 0; k < NGEMM; k++) {
    Gemm& g = P->g[k];
    g.M = shapes[k][0], g.N = shapes[k][1], g.K = shapes[k][2];
    if (P->lt) {
      if (int r = make_gemm(c, *P, g, g.M, g.N, g.K)) return r;
    } else if (int r = plan_hand(c, g, cmax)) {
      return r;
    }
  }
  if (P->lt) {
    HIPCHK(c, dalloc(P->ws, kWorkspace));
    // keep each product's fastest algorithm (timed on zero data of the real shapes)
    if (int r = tune_gemm(c, *P, P->g[G1], P->a[0].p, P->rho3.p, st)) return r;
    if (int r = tune_gemm(c, *P, P->g[G2], P->a[1].p, P->t13.p, st)) return r;
    if (int r = tune_gemm(c, *P, P->g[G3], P->a[2].p, P->p3.p, st)) return r;
    if (int r = tune_gemm(c, *P, P->g[G4_PHI], P->a[3].p, P->r3.p, st)) return r;
    if (int r = tune_gemm(c, *P, P->g[G4], P->a[3].p, P->r3.p + 3 * B, st)) return r;
    if (int r = tune_gemm(c, *P, P->g[G5], P->a[4].p, P->r3.p, st)) return r;
  }
  HIPCHK(c, hipStreamSynchronize(st));
  size_t bytes = 0;
  for (auto& x : P->a) bytes += x.n * 2;
  bytes += (P->rho3.n + P->t13.n + P->p3.n + P->r3.n) * 2 + P->c.n * 4 + P->ws.n;
  m.nbytes = bytes;
  m.ready = true;
  return 0;
}

// rho (d.rho) -> d.phi (if want_phi) and d.field through G1..G5.
hipError_t rp_launch_poisson_mfma(const RpDev& d, const RpMfmaDev& mv, hipStream_t s, int want_phi) {
  const Plan& P = *static_cast<const Plan*>(mv.plan);
  const int nx = d.nbx, ny = d.nby;
  const size_t B = (size_t)nx * ny;
  int ns = 1;
  hipLaunchKernelGGL(k_rho_parts, dim3(MEAN_PARTS), dim3(256), 0, s, (const float*)d.rho, B, P.sums.p);
  hipLaunchKernelGGL(k_rho_mean, dim3(1), dim3(256), 0, s, (const double*)P.sums.p, B, P.sums.p + MEAN_PARTS);
  hipLaunchKernelGGL(k_split_rho, grid_of(B), dim3(256), 0, s, (const float*)d.rho,
                     (const double*)(P.sums.p + MEAN_PARTS), (rp_bf16*)P.rho3.p, nx, ny);
  // G1: T1t[u][j] = Cx[u][i] . rho[j][i]
  hipError_t e = product(P, P.g[G1], P.a[0].p, P.rho3.p, P.c.p, s, &ns);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_split_c, grid_of(B), dim3(256), 0, s, (const float*)P.c.p, (rp_bf16*)P.t13.p, nx, ny, 0,
                     (const float*)nullptr, (const float*)nullptr, 0.0f, ns);
  // G2: P[v][u] = Cy[v][j] . T1t[u][j], scaled
  if ((e = product(P, P.g[G2], P.a[1].p, P.t13.p, P.c.p, s, &ns)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_split_c, grid_of(B), dim3(256), 0, s, (const float*)P.c.p, (rp_bf16*)P.p3.p, ny, nx, 1,
                     (const float*)P.wx2.p, (const float*)P.wy2.p, (float)(4.0 / nx / ny), ns);
  // G3: [R1t; R2t][i][v] = [Cxt; wx Sxt][i][u] . P[v][u]
  if ((e = product(P, P.g[G3], P.a[2].p, P.p3.p, P.c.p, s, &ns)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_split_c, grid_of(2 * B), dim3(256), 0, s, (const float*)P.c.p, (rp_bf16*)P.r3.p, 2 * nx, ny,
                     0, (const float*)nullptr, (const float*)nullptr, 0.0f, ns);
  // G4: [phi | Ex][j][.] = Cyt[j][v] . [R1t; R2t][.][v]  (Ex only without phi)
  if (want_phi) {
    if ((e = product(P, P.g[G4_PHI], P.a[3].p, P.r3.p, P.c.p, s, &ns)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_phi_ex, grid_of(2 * B), dim3(256), 0, s, (const float*)P.c.p, d.phi, d.field, nx, ny,
                       2 * nx, 0, ns);
  } else {
    if ((e = product(P, P.g[G4], P.a[3].p, P.r3.p + 3 * B, P.c.p, s, &ns)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_phi_ex, grid_of(B), dim3(256), 0, s, (const float*)P.c.p, d.phi, d.field, nx, ny, nx, nx,
                       ns);
  }
  // G5: Ey[j][i] = (wy Syt)[j][v] . R1t[i][v]
  if ((e = product(P, P.g[G5], P.a[4].p, P.r3.p, P.c.p, s, &ns)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_ey, grid_of(B), dim3(256), 0, s, (const float*)P.c.p, d.field, B, ns);
  return hipGetLastError();
}
