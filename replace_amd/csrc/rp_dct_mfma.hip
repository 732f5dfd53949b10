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
//   G2  P[v][u]            = Cy[v][j]            . T1t[u][j]     (spectral scaling in the epilogue)
//   G3  [R1t; R2t][i][v]   = [Cxt; wx Sxt][i][u] . P[v][u]
//   G4  [phi | Ex][j][.]   = Cyt[j][v]           . [R1t; R2t][.][v]
//   G5  Ey[j][i]           = (wy Syt)[j][v]      . R1t[i][v]
// 14 N^3 multiply-adds per evaluation at nx = ny = N (12 N^3 without phi).
// Each product of split operands is hi.hi + hi.lo + lo.hi (the lo.lo term is
// below FP32 rounding of the accumulation), i.e. three bf16 MFMAs.
#include <cmath>

#include "rp_ctx.h"

typedef __bf16 rp_bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

namespace {

constexpr int GT = 128;      // block tile (M and N)
constexpr int GK = 32;       // k step staged in LDS
constexpr int GKP = GK + 8;  // padded LDS row (bank spread for 16-byte fragment reads)

enum { EPI_SPLIT = 0, EPI_SPECTRAL = 1, EPI_PHI_EX = 2, EPI_EY = 3 };

struct SplitMat {
  const rp_bf16* hi;
  const rp_bf16* lo;
  int ld;
};

struct GemmArgs {
  SplitMat a, b;
  int M, N, K;
  // EPI_SPLIT / EPI_SPECTRAL output
  rp_bf16* ohi;
  rp_bf16* olo;
  int ldc;
  // EPI_SPECTRAL
  const float* wx2;  // (pi u / nx)^2
  const float* wy2;  // (pi v / ny * asp)^2
  float wscale;      // 4 / (nx ny)
  // EPI_PHI_EX / EPI_EY
  float* phi;
  float2* field;
  int nx, col0;  // col0: first output column's index in [phi | Ex]
};

__device__ __forceinline__ void split_bf16(float v, rp_bf16& h, rp_bf16& l) {
  h = (rp_bf16)v;
  l = (rp_bf16)(v - (float)h);
}

template <int EPI>
__device__ __forceinline__ void epilogue(const GemmArgs& g, int m, int n, float v) {
  if (EPI == EPI_SPLIT) {
    rp_bf16 h, l;
    split_bf16(v, h, l);
    g.ohi[(size_t)m * g.ldc + n] = h;
    g.olo[(size_t)m * g.ldc + n] = l;
  } else if (EPI == EPI_SPECTRAL) {  // m = v, n = u
    float s = 0.0f;
    if (m != 0 || n != 0) {
      float w = g.wscale;
      if (n == 0) w *= 0.5f;
      if (m == 0) w *= 0.5f;
      s = w / (g.wx2[n] + g.wy2[m]);
    }
    rp_bf16 h, l;
    split_bf16(v * s, h, l);
    g.ohi[(size_t)m * g.ldc + n] = h;
    g.olo[(size_t)m * g.ldc + n] = l;
  } else if (EPI == EPI_PHI_EX) {  // m = j, column col0 + n of [phi | Ex]
    int c = g.col0 + n;
    if (c < g.nx)
      g.phi[(size_t)m * g.nx + c] = v;
    else
      g.field[(size_t)m * g.nx + (c - g.nx)].x = v;
  } else {  // EPI_EY
    g.field[(size_t)m * g.nx + n].y = v;
  }
}

// One 128 x 128 output tile per 256-thread workgroup; each of the four waves
// owns a 64 x 64 quarter as 2 x 2 tiles of v_mfma_f32_32x32x16_bf16.
// Fragment layout of the 32x32x16 instruction: lane l supplies row (l % 32)
// of A and column (l % 32) of B with k = 8 (l / 32) .. +7; accumulator
// element e of lane l is row 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32.
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_split(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) rp_bf16 As[2][GT][GKP];
  __shared__ __attribute__((aligned(16))) rp_bf16 Bs[2][GT][GKP];
  const int tiles_n = g.N / GT;
  const int nt = tiles_n * (g.M / GT);
  // XCD-aware order: hardware block b runs on XCD b % 8; give each XCD a
  // contiguous run of tiles (neighbouring tiles share A rows in L2)
  int b = blockIdx.x;
  if (nt % 8 == 0) b = (b % 8) * (nt / 8) + b / 8;
  const int tm = b / tiles_n, tn = b % tiles_n;
  const int bm = tm * GT, bn = tn * GT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

  floatx16 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

  // staging: 128 rows x 32 k = 512 chunks of 8 bf16 per matrix and part
  uint4 ra[2][2], rb[2][2];  // [part][chunk]
  auto gload = [&](int k0) {
    for (int c = 0; c < 2; c++) {
      int ch = tid + c * 256, row = ch >> 2, kc = (ch & 3) * 8;
      size_t oa = (size_t)(bm + row) * g.a.ld + k0 + kc;
      size_t ob = (size_t)(bn + row) * g.b.ld + k0 + kc;
      ra[0][c] = *(const uint4*)(g.a.hi + oa);
      ra[1][c] = *(const uint4*)(g.a.lo + oa);
      rb[0][c] = *(const uint4*)(g.b.hi + ob);
      rb[1][c] = *(const uint4*)(g.b.lo + ob);
    }
  };
  auto lstore = [&]() {
    for (int c = 0; c < 2; c++) {
      int ch = tid + c * 256, row = ch >> 2, kc = (ch & 3) * 8;
      for (int p = 0; p < 2; p++) {
        *(uint4*)&As[p][row][kc] = ra[p][c];
        *(uint4*)&Bs[p][row][kc] = rb[p][c];
      }
    }
  };
  gload(0);
  lstore();
  __syncthreads();
  const int fr = lane & 31, fk = (lane >> 5) * 8;
  for (int k0 = 0; k0 < g.K; k0 += GK) {
    const bool more = k0 + GK < g.K;
    if (more) gload(k0 + GK);
    for (int kk = 0; kk < GK; kk += 16) {
      bf16x8 ah[2], al[2], bh[2], bl[2];
      for (int i = 0; i < 2; i++) {
        ah[i] = *(const bf16x8*)&As[0][wm + i * 32 + fr][kk + fk];
        al[i] = *(const bf16x8*)&As[1][wm + i * 32 + fr][kk + fk];
        bh[i] = *(const bf16x8*)&Bs[0][wn + i * 32 + fr][kk + fk];
        bl[i] = *(const bf16x8*)&Bs[1][wn + i * 32 + fr][kk + fk];
      }
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
    if (more) {
      lstore();
      __syncthreads();
    }
  }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int e = 0; e < 16; e++) {
        int m = bm + wm + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        int n = bn + wn + j * 32 + (lane & 31);
        epilogue<EPI>(g, m, n, acc[i][j][e]);
      }
}

// Mean of rho (fixed-order FP64 partials, then one block).  Subtracting a
// constant from rho changes only the u = v = 0 coefficient, which the
// spectral step discards, so the transform of (rho - mean) is exact; it keeps
// the 16-bit rounding of the large mean out of the low modes that the
// spectral division amplifies.
constexpr int MEAN_PARTS = 256;

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
  double s = block_sum(parts[threadIdx.x], sh);
  if (threadIdx.x == 0) mean[0] = s / (double)n;
}

// rho (FP32, [ny][nx]) - mean -> split pair
__global__ void k_split(const float* __restrict__ x, const double* __restrict__ mean, rp_bf16* __restrict__ hi,
                        rp_bf16* __restrict__ lo, size_t n) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    rp_bf16 h, l;
    split_bf16((float)((double)x[k] - mean[0]), h, l);
    hi[k] = h;
    lo[k] = l;
  }
}

template <int EPI>
hipError_t gemm(const GemmArgs& g, hipStream_t s) {
  int nt = (g.M / GT) * (g.N / GT);
  hipLaunchKernelGGL(k_gemm_split<EPI>, dim3(nt), dim3(256), 0, s, g);
  return hipGetLastError();
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

void split_host(const std::vector<double>& v, std::vector<uint16_t>& hi, std::vector<uint16_t>& lo) {
  hi.resize(v.size());
  lo.resize(v.size());
  for (size_t k = 0; k < v.size(); k++) {
    float f = (float)v[k];
    hi[k] = bf16_bits(f);
    lo[k] = bf16_bits(f - bf16_float(hi[k]));
  }
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

}  // namespace

// Transform matrices of the dense split-BF16 Poisson solve (once per set-up).
int rp_mfma_setup(rp_ctx* c) {
  const int nx = c->nbx, ny = c->nby;
  if (nx % GT || ny % GT)
    return fail(c, RP_ERR_ARG, "matrix-core Poisson needs bin counts that are multiples of %d (got %dx%d)", GT,
                nx, ny);
  hipStream_t st = c->stream;
  RpMfma& m = c->mfma;
  const double asp = (double)c->bsy / (double)c->bsx;
  auto put = [&](const std::vector<double>& v, DevBuf<uint16_t>& hi, DevBuf<uint16_t>& lo) -> int {
    std::vector<uint16_t> h, l;
    split_host(v, h, l);
    HIPCHK(c, dupload(hi, h, st));
    HIPCHK(c, dupload(lo, l, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return 0;
  };
  std::vector<double> M((size_t)2 * nx * nx);
  // Cx[u][i]
  M.resize((size_t)nx * nx);
  for (int u = 0; u < nx; u++)
    for (int i = 0; i < nx; i++) M[(size_t)u * nx + i] = cospi_frac((long long)u * (2 * i + 1), 2LL * nx);
  if (int r = put(M, m.cx_hi, m.cx_lo)) return r;
  // [Cxt; wx Sxt][i][u]
  M.resize((size_t)2 * nx * nx);
  for (int i = 0; i < nx; i++)
    for (int u = 0; u < nx; u++) {
      M[(size_t)i * nx + u] = cospi_frac((long long)u * (2 * i + 1), 2LL * nx);
      M[(size_t)(nx + i) * nx + u] = (M_PI * u / nx) * sinpi_frac((long long)u * (2 * i + 1), 2LL * nx);
    }
  if (int r = put(M, m.cxs_hi, m.cxs_lo)) return r;
  // Cy[v][j], Cyt[j][v], (wy Syt)[j][v]
  M.resize((size_t)ny * ny);
  for (int v = 0; v < ny; v++)
    for (int j = 0; j < ny; j++) M[(size_t)v * ny + j] = cospi_frac((long long)v * (2 * j + 1), 2LL * ny);
  if (int r = put(M, m.cy_hi, m.cy_lo)) return r;
  for (int j = 0; j < ny; j++)
    for (int v = 0; v < ny; v++) M[(size_t)j * ny + v] = cospi_frac((long long)v * (2 * j + 1), 2LL * ny);
  if (int r = put(M, m.cyt_hi, m.cyt_lo)) return r;
  for (int j = 0; j < ny; j++)
    for (int v = 0; v < ny; v++)
      M[(size_t)j * ny + v] = (M_PI * v / ny * asp) * sinpi_frac((long long)v * (2 * j + 1), 2LL * ny);
  if (int r = put(M, m.sy_hi, m.sy_lo)) return r;
  std::vector<float> wx2(nx), wy2(ny);
  for (int u = 0; u < nx; u++) wx2[u] = (float)((M_PI * u / nx) * (M_PI * u / nx));
  for (int v = 0; v < ny; v++) wy2[v] = (float)((M_PI * v / ny * asp) * (M_PI * v / ny * asp));
  HIPCHK(c, dupload(m.wx2, wx2, st));
  HIPCHK(c, dupload(m.wy2, wy2, st));
  const size_t B = (size_t)nx * ny;
  HIPCHK(c, dalloc(m.rho_hi, B));
  HIPCHK(c, dalloc(m.rho_lo, B));
  HIPCHK(c, dalloc(m.t1_hi, B));
  HIPCHK(c, dalloc(m.t1_lo, B));
  HIPCHK(c, dalloc(m.p_hi, B));
  HIPCHK(c, dalloc(m.p_lo, B));
  HIPCHK(c, dalloc(m.r_hi, 2 * B));
  HIPCHK(c, dalloc(m.r_lo, 2 * B));
  HIPCHK(c, dalloc(m.sums, MEAN_PARTS + 1));
  HIPCHK(c, hipStreamSynchronize(st));
  m.ready = true;
  return 0;
}

// rho (d.rho) -> d.phi (if want_phi) and d.field through G1..G5.
hipError_t rp_launch_poisson_mfma(const RpDev& d, const RpMfmaDev& m, hipStream_t s, int want_phi) {
  const int nx = d.nbx, ny = d.nby;
  const size_t B = (size_t)nx * ny;
  hipLaunchKernelGGL(k_rho_parts, dim3(MEAN_PARTS), dim3(256), 0, s, (const float*)d.rho, B, m.sums);
  hipLaunchKernelGGL(k_rho_mean, dim3(1), dim3(MEAN_PARTS), 0, s, (const double*)m.sums, B, m.sums + MEAN_PARTS);
  hipLaunchKernelGGL(k_split, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, (const float*)d.rho,
                     (const double*)(m.sums + MEAN_PARTS), m.rho_hi, m.rho_lo, B);
  GemmArgs g{};
  g.wx2 = m.wx2;
  g.wy2 = m.wy2;
  g.wscale = (float)(4.0 / nx / ny);
  g.phi = d.phi;
  g.field = d.field;
  g.nx = nx;
  // G1: T1t[u][j] = Cx[u][i] . rho[j][i]
  g.a = {m.cx_hi, m.cx_lo, nx};
  g.b = {m.rho_hi, m.rho_lo, nx};
  g.M = nx, g.N = ny, g.K = nx;
  g.ohi = m.t1_hi, g.olo = m.t1_lo, g.ldc = ny;
  hipError_t e = gemm<EPI_SPLIT>(g, s);
  if (e != hipSuccess) return e;
  // G2: P[v][u] = Cy[v][j] . T1t[u][j], scaled
  g.a = {m.cy_hi, m.cy_lo, ny};
  g.b = {m.t1_hi, m.t1_lo, ny};
  g.M = ny, g.N = nx, g.K = ny;
  g.ohi = m.p_hi, g.olo = m.p_lo, g.ldc = nx;
  if ((e = gemm<EPI_SPECTRAL>(g, s)) != hipSuccess) return e;
  // G3: [R1t; R2t][i][v] = [Cxt; wx Sxt][i][u] . P[v][u]
  g.a = {m.cxs_hi, m.cxs_lo, nx};
  g.b = {m.p_hi, m.p_lo, nx};
  g.M = 2 * nx, g.N = ny, g.K = nx;
  g.ohi = m.r_hi, g.olo = m.r_lo, g.ldc = ny;
  if ((e = gemm<EPI_SPLIT>(g, s)) != hipSuccess) return e;
  // G4: [phi | Ex][j][.] = Cyt[j][v] . [R1t; R2t][.][v]  (Ex only without phi)
  g.a = {m.cyt_hi, m.cyt_lo, ny};
  if (want_phi) {
    g.b = {m.r_hi, m.r_lo, ny};
    g.N = 2 * nx;
    g.col0 = 0;
  } else {
    g.b = {m.r_hi + B, m.r_lo + B, ny};
    g.N = nx;
    g.col0 = nx;
  }
  g.M = ny, g.K = ny;
  if ((e = gemm<EPI_PHI_EX>(g, s)) != hipSuccess) return e;
  // G5: Ey[j][i] = (wy Syt)[j][v] . R1t[i][v]
  g.a = {m.sy_hi, m.sy_lo, ny};
  g.b = {m.r_hi, m.r_lo, ny};
  g.M = ny, g.N = nx, g.K = ny;
  return gemm<EPI_EY>(g, s);
}
