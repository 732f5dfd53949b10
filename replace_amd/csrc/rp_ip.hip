// rp_ip.hip -- InitialPlace's B2B linear systems assembled on the device.
//
// Reference: InitialPlace::updatePinInfo (src/initialPlace.cpp:128-182) and
// createSparseMatrix (src/initialPlace.cpp:185-378), whose triplet lists go
// through Eigen's SparseMatrix::setFromTriplets (src/initialPlace.cpp:376-377).
//
//   1. Per net, one thread finds the pins that end as the x/y min and max of
//      updatePinInfo's sequential scan (the first pin reaching a strict new
//      extreme keeps the flag).
//   2. Per pin a1, one thread walks the reference's pair loop (a2 <= a1,
//      same-instance pairs skipped) over the pairs that touch an extreme pin,
//      emitting the Laplacian triplets and the right-hand-side terms at
//      offsets from an exclusive scan of per-pin counts; pins are net-major,
//      so this is the reference's order.
// The CSR is then what setFromTriplets builds: a stable radix sort of the
// triplets on (row, col) keeps equal entries in triplet order, and every
// (row, col) run is summed sequentially from 0.0f.  The right-hand side is
// summed per row in the same order the reference's `b[k] += ...` adds it.
// Integer keys and in-order float sums make the result bit-identical to a
// sequential host assembly, independent of launch shape.
#include <chrono>
#include <memory>
#include <cstdlib>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "rp_ctx.h"

namespace {

struct IpView {
  int n_net, n, n_pin;
  const int* __restrict__ net_start;
  const int* __restrict__ pin_inst;
  const int* __restrict__ pin_k;   // movable index of the pin's instance, -1 for fixed / BTerm
  const int* __restrict__ pin_cx;
  const int* __restrict__ pin_cy;
  const int* __restrict__ icx;     // integer centre of movable instance k
  const int* __restrict__ icy;
  int min_diff, max_fanout;
  float weight_scale;
};

struct IpOut {
  const int* off;  // 4 x (n_pin + 1) exclusive offsets: x triplets, y triplets, x rhs, y rhs
  unsigned long long* tk[2];
  float* tv[2];
  unsigned* rk[2];
  float* rv[2];
  int cb;    // column bits of the packed key
};

struct Emit {
  unsigned long long* tk;
  float* tv;
  unsigned* rk;
  float* rv;
  int nt, nr, cb;
  __device__ void trip(int r, int c, float v) {
    tk[nt] = ((unsigned long long)(unsigned)r << cb) | (unsigned)c;
    tv[nt] = v;
    nt++;
  }
  __device__ void rhs(int r, float v) {
    rk[nr] = (unsigned)r;
    rv[nr] = v;
    nr++;
  }
};

// One B2B pair on one axis (src/initialPlace.cpp:242-290 for x, :300-348 for y).
template <bool EMIT>
__device__ __forceinline__ void ip_pair(int p1, int p2, int k1, int k2, const int* ic, float netWeight, float minDiff,
                                        int min_diff, int& nt, int& nr, Emit& em) {
  int diff = abs(p1 - p2);
  float w = diff > min_diff ? netWeight / (float)diff : netWeight / minDiff;
  if (k1 >= 0 && k2 >= 0) {
    if (EMIT) {
      em.trip(k1, k1, w);
      em.trip(k2, k2, w);
      em.trip(k1, k2, -w);
      em.trip(k2, k1, -w);
      em.rhs(k1, -w * (float)((p1 - ic[k1]) - (p2 - ic[k2])));
      em.rhs(k2, -w * (float)((p2 - ic[k2]) - (p1 - ic[k1])));
    }
    nt += 4;
    nr += 2;
  } else if (k1 < 0 && k2 >= 0) {
    if (EMIT) {
      em.trip(k2, k2, w);
      em.rhs(k2, w * (float)(p1 - (p2 - ic[k2])));
    }
    nt += 1;
    nr += 1;
  } else if (k1 >= 0 && k2 < 0) {
    if (EMIT) {
      em.trip(k1, k1, w);
      em.rhs(k1, w * (float)(p2 - (p1 - ic[k1])));
    }
    nt += 1;
    nr += 1;
  }
}

// updatePinInfo for net e: the pins that end as the x/y min and max of the
// reference's sequential scan (the first pin reaching a strict new extreme
// keeps the flag); all -1 for nets the B2B model skips.  The loads of several
// pins are independent, so the scan is latency-light even for long nets.
__global__ void __launch_bounds__(RP_BLOCK) k_ip_ext(IpView v, int4* __restrict__ ext) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= v.n_net) return;
  const int s = v.net_start[e], t = v.net_start[e + 1], deg = t - s;
  if (deg <= 1 || deg >= v.max_fanout) {
    ext[e] = make_int4(-1, -1, -1, -1);
    return;
  }
  int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
  int mnx = -1, mxx = -1, mny = -1, mxy = -1;
#pragma unroll 4
  for (int q = s; q < t; q++) {
    const int x = v.pin_cx[q], y = v.pin_cy[q];
    if (lx > x) { lx = x; mnx = q; }
    if (ux < x) { ux = x; mxx = q; }
    if (ly > y) { ly = y; mny = q; }
    if (uy < y) { uy = y; mxy = q; }
  }
  ext[e] = make_int4(mnx, mxx, mny, mxy);
}

// The reference's pair loop (a1 over the net's pins, a2 <= a1, same-instance
// pairs skipped) for one a1 = pin q1 of net [s, t), restricted to the pairs
// that touch an extreme pin of the axis, partners visited in ascending order.
// Visiting the pins in order gives the reference's triplet order per axis.
template <bool EMIT>
__device__ void ip_pin(const IpView& v, int q1, int s, int t, int4 ext, int cnt[4], Emit* ex, Emit* ey) {
  cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
  const int mnx = ext.x, mxx = ext.y, mny = ext.z, mxy = ext.w;
  if (mnx < 0) return;  // net skipped
  const float netWeight = v.weight_scale / (float)(t - s - 1);
  const float minDiff = (float)v.min_diff;
  Emit dummy{};
  Emit& emx = EMIT ? *ex : dummy;
  Emit& emy = EMIT ? *ey : dummy;
  const int i1 = v.pin_inst[q1], k1 = v.pin_k[q1];
  const int x1 = v.pin_cx[q1], y1 = v.pin_cy[q1];
  const bool xe1 = q1 == mnx || q1 == mxx, ye1 = q1 == mny || q1 == mxy;
  if (xe1 || ye1) {
    for (int q2 = s; q2 <= q1; q2++) {
      const int i2 = v.pin_inst[q2];
      if (i1 == i2) continue;
      const int k2 = v.pin_k[q2];
      if (xe1 || q2 == mnx || q2 == mxx)
        ip_pair<EMIT>(x1, v.pin_cx[q2], k1, k2, v.icx, netWeight, minDiff, v.min_diff, cnt[0], cnt[2], emx);
      if (ye1 || q2 == mny || q2 == mxy)
        ip_pair<EMIT>(y1, v.pin_cy[q2], k1, k2, v.icy, netWeight, minDiff, v.min_diff, cnt[1], cnt[3], emy);
    }
  } else {
    // only extreme partners with q2 <= q1, ascending; x and y lists are independent
    const int ax = min(mnx, mxx), bx = max(mnx, mxx);
    const int ay = min(mny, mxy), by = max(mny, mxy);
    const int xs[2] = {ax, bx == ax ? -1 : bx};
    const int ys[2] = {ay, by == ay ? -1 : by};
    for (int j = 0; j < 2; j++) {
      const int q2 = xs[j];
      if (q2 < 0 || q2 > q1) continue;
      if (i1 == v.pin_inst[q2]) continue;
      ip_pair<EMIT>(x1, v.pin_cx[q2], k1, v.pin_k[q2], v.icx, netWeight, minDiff, v.min_diff, cnt[0], cnt[2], emx);
    }
    for (int j = 0; j < 2; j++) {
      const int q2 = ys[j];
      if (q2 < 0 || q2 > q1) continue;
      if (i1 == v.pin_inst[q2]) continue;
      ip_pair<EMIT>(y1, v.pin_cy[q2], k1, v.pin_k[q2], v.icy, netWeight, minDiff, v.min_diff, cnt[1], cnt[3], emy);
    }
  }
}

// per pin: triplet and right-hand-side counts of both axes
__global__ void __launch_bounds__(RP_BLOCK) k_ip_count(IpView v, const int* __restrict__ pin_net,
                                                       const int4* __restrict__ ext, int* cnt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = v.n_pin + 1;
  if (q == v.n_pin) {
    for (int a = 0; a < 4; a++) cnt[a * stride + q] = 0;
  }
  if (q >= v.n_pin) return;
  const int e = pin_net[q];
  int c[4];
  ip_pin<false>(v, q, v.net_start[e], v.net_start[e + 1], ext[e], c, nullptr, nullptr);
  for (int a = 0; a < 4; a++) cnt[a * stride + q] = c[a];
}

__global__ void __launch_bounds__(RP_BLOCK) k_ip_emit(IpView v, const int* __restrict__ pin_net,
                                                      const int4* __restrict__ ext, IpOut o) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= v.n_pin) return;
  const int stride = v.n_pin + 1;
  Emit ex{o.tk[0], o.tv[0], o.rk[0], o.rv[0], o.off[0 * stride + q], o.off[2 * stride + q], o.cb};
  Emit ey{o.tk[1], o.tv[1], o.rk[1], o.rv[1], o.off[1 * stride + q], o.off[3 * stride + q], o.cb};
  const int e = pin_net[q];
  int c[4];
  ip_pin<true>(v, q, v.net_start[e], v.net_start[e + 1], ext[e], c, &ex, &ey);
}

// run heads of a sorted key array
template <typename K>
__global__ void k_ip_heads(const K* __restrict__ key, int m, int* __restrict__ flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// one (row, col) run -> one CSR entry, summed in triplet order from 0.0f
// Sum of the run of equal keys starting at i, in order from 0.0f: eight
// keys and values are loaded before they are added (a long run, such as the
// diagonal entry of a cell on many nets, is walked eight entries a round).
template <typename K>
__device__ __forceinline__ float run_sum(const K* __restrict__ key, const float* __restrict__ val, int m, int i,
                                         K k) {
  float s = 0.0f;
  for (int j = i; j < m; j += 8) {
    K kk[8];
    float vv[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int q = min(j + u, m - 1);
      kk[u] = key[q];
      vv[u] = val[q];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      if (j + u >= m || kk[u] != k) return s;
      s += vv[u];
    }
  }
  return s;
}

__global__ void k_ip_segsum(const unsigned long long* __restrict__ key, const float* __restrict__ val, int m,
                            const int* __restrict__ flag, const int* __restrict__ seg, int cb,
                            int* __restrict__ col, float* __restrict__ out, int* __restrict__ srow) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !flag[i]) return;
  const unsigned long long k = key[i];
  const float s = run_sum(key, val, m, i, k);
  int at = seg[i];
  col[at] = (int)(k & ((1ull << cb) - 1));
  srow[at] = (int)(k >> cb);
  out[at] = s;
}

// rowp[r] = first entry whose row >= r
__global__ void k_ip_rowp(const int* __restrict__ srow, const int* __restrict__ flag, const int* __restrict__ seg,
                          int m, int n, int* __restrict__ rowp) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  int nseg = m > 0 ? seg[m - 1] + flag[m - 1] : 0;
  int lo = 0, hi = nseg;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (srow[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  rowp[r] = lo;
}

__global__ void k_ip_zero(float* __restrict__ b, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = 0.0f;
}

// right-hand side: per row, the terms in triplet order, from 0.0f
__global__ void k_ip_rhs(const unsigned* __restrict__ key, const float* __restrict__ val, int m,
                         float* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || (i > 0 && key[i] == key[i - 1])) return;
  const unsigned k = key[i];
  b[k] = run_sum(key, val, m, i, k);
}

inline dim3 grid_of(long n) { return dim3((unsigned)std::max<long>(1, (n + RP_BLOCK - 1) / RP_BLOCK)); }

template <typename T>
hipError_t ensure(DevBuf<T>& b, size_t n) {
  if (b.n >= n && b.p) return hipSuccess;
  return dalloc(b, n + n / 4 + 64);
}

int bits_for(int n) {
  int b = 1;
  while ((1ll << b) < (long long)n) b++;
  return b;
}

}  // namespace

int rp_ip_asm_init(rp_ctx* c, RpIpAsm& a) {
  hipStream_t st = c->stream;
  const int n = (int)c->place.size();
  a.n = n;
  a.n_net = c->n_net;
  a.cb = bits_for(std::max(2, n));
  // rows and columns in locality order: row pos[k] holds movable instance k
  const std::vector<int>& order = rp_movable_order(c);
  a.pos.assign(n, 0);
  for (int j = 0; j < n; j++) a.pos[order[j]] = j;
  std::vector<int>& pk = a.h_pin_k;
  pk.assign(c->n_pin, -1);
  for (int q = 0; q < c->n_pin; q++) {
    int i = c->pin_inst[q];
    int k = i >= 0 ? c->inst2place[i] : -1;
    pk[q] = k >= 0 ? a.pos[k] : -1;
  }
  std::vector<int> pin_net(c->n_pin);
  for (int e = 0; e < c->n_net; e++)
    for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) pin_net[q] = e;
  HIPCHK(c, dupload(a.pin_net, pin_net, st));
  HIPCHK(c, dalloc(a.ext, std::max(1, c->n_net)));
  HIPCHK(c, dupload(a.net_start, c->net_start, st));
  HIPCHK(c, dupload(a.pin_inst, c->pin_inst, st));
  HIPCHK(c, dupload(a.pin_k, pk, st));
  HIPCHK(c, dalloc(a.pin_cx, c->n_pin));
  HIPCHK(c, dalloc(a.pin_cy, c->n_pin));
  HIPCHK(c, dalloc(a.icx, n));
  HIPCHK(c, dalloc(a.icy, n));
  HIPCHK(c, dalloc(a.cnt, 4 * ((size_t)c->n_pin + 1)));
  HIPCHK(c, dalloc(a.off, 4 * ((size_t)c->n_pin + 1)));
  for (int ax = 0; ax < 2; ax++) {
    HIPCHK(c, dalloc(a.rowp[ax], (size_t)n + 1));
    HIPCHK(c, dalloc(a.b[ax], n));
  }
  if (!a.h_tot) HIPCHK(c, rp_pinned_alloc((void**)&a.h_tot, 4 * sizeof(int)));
  HIPCHK(c, hipStreamSynchronize(st));
  return 0;
}

// Assemble both axes from the device's current pin and instance centres
// (a.pin_cx / a.pin_cy / a.icx / a.icy).
// First half: per-pin counts and their scans; the four totals are copied to
// a.h_tot, which the caller waits for before the second half.
int rp_ip_assemble_a(rp_ctx* c, RpIpAsm& a, hipStream_t st) {
  const int n = a.n, nn = a.n_net, np = c->n_pin;
  const size_t stride = (size_t)np + 1;
  IpView v{nn, n, np, a.net_start.p, a.pin_inst.p, a.pin_k.p, a.pin_cx.p, a.pin_cy.p, a.icx.p, a.icy.p,
           c->prm.initial_place_min_diff_length, c->prm.initial_place_max_fanout,
           c->prm.initial_place_net_weight_scale};
  hipLaunchKernelGGL(k_ip_ext, grid_of(nn), dim3(RP_BLOCK), 0, st, v, a.ext.p);
  hipLaunchKernelGGL(k_ip_count, grid_of((long)np + 1), dim3(RP_BLOCK), 0, st, v, (const int*)a.pin_net.p,
                     (const int4*)a.ext.p, a.cnt.p);
  HIPCHK(c, hipGetLastError());
  // exclusive scans of the four count lists (length np + 1: the last entry is the total)
  size_t need = 0;
  HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, need, a.cnt.p, a.off.p, (int)stride, st));
  HIPCHK(c, ensure(a.temp, need));
  for (int q = 0; q < 4; q++) {
    size_t tb = a.temp.n;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(a.temp.p, tb, a.cnt.p + q * stride, a.off.p + q * stride,
                                               (int)stride, st));
  }
  for (int q = 0; q < 4; q++)
    HIPCHK(c, hipMemcpyAsync(a.h_tot + q, a.off.p + q * stride + np, sizeof(int), hipMemcpyDeviceToHost, st));
  return 0;
}

// Second half: triplets, CSR and right-hand sides of both axes.
int rp_ip_assemble_b(rp_ctx* c, RpIpAsm& a, hipStream_t st) {
  const int n = a.n, nn = a.n_net, np = c->n_pin;
  IpView v{nn, n, np, a.net_start.p, a.pin_inst.p, a.pin_k.p, a.pin_cx.p, a.pin_cy.p, a.icx.p, a.icy.p,
           c->prm.initial_place_min_diff_length, c->prm.initial_place_max_fanout,
           c->prm.initial_place_net_weight_scale};
  int mt[2] = {a.h_tot[0], a.h_tot[1]}, mr[2] = {a.h_tot[2], a.h_tot[3]};
  const auto te = std::chrono::steady_clock::now();
  for (int ax = 0; ax < 2; ax++) {
    if (mt[ax] < 0 || mr[ax] < 0) return fail(c, RP_ERR_NOMEM, "initial place: B2B system exceeds 2^31 entries");
    HIPCHK(c, ensure(a.tk[ax], mt[ax]));
    HIPCHK(c, ensure(a.tk2[ax], mt[ax]));
    HIPCHK(c, ensure(a.tv[ax], mt[ax]));
    HIPCHK(c, ensure(a.tv2[ax], mt[ax]));
    HIPCHK(c, ensure(a.rk[ax], mr[ax]));
    HIPCHK(c, ensure(a.rk2[ax], mr[ax]));
    HIPCHK(c, ensure(a.rv[ax], mr[ax]));
    HIPCHK(c, ensure(a.rv2[ax], mr[ax]));
    HIPCHK(c, ensure(a.col[ax], mt[ax]));
    HIPCHK(c, ensure(a.val[ax], mt[ax]));
  }
  const int mmax = std::max(mt[0], mt[1]);
  HIPCHK(c, ensure(a.flag, mmax));
  HIPCHK(c, ensure(a.seg, mmax));
  HIPCHK(c, ensure(a.srow, mmax));
  // the sort / scan temporary storage of both axes, sized before anything is
  // queued: growing it between the axes would free a buffer that the first
  // axis' queued sorts still use
  size_t tneed = 0;
  for (int ax = 0; ax < 2; ax++) {
    size_t t1 = 0, t2 = 0, t3 = 0;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, t1, a.tk[ax].p, a.tk2[ax].p, a.tv[ax].p, a.tv2[ax].p,
                                                 std::max(1, mt[ax]), 0, 2 * a.cb, st));
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, t2, a.rk[ax].p, a.rk2[ax].p, a.rv[ax].p, a.rv2[ax].p,
                                                 std::max(1, mr[ax]), 0, a.cb, st));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, t3, a.flag.p, a.seg.p, std::max(1, mt[ax]), st));
    tneed = std::max(tneed, std::max(t1, std::max(t2, t3)));
  }
  HIPCHK(c, ensure(a.temp, tneed));
  {
    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - te).count();
    if (t > 10.0 && getenv("RP_IP_TIMING")) fprintf(stderr, "[rp] slow assembly buffers: %.1f ms (%d rows)\n", t, n);
  }
  IpOut o{a.off.p, {a.tk[0].p, a.tk[1].p}, {a.tv[0].p, a.tv[1].p}, {a.rk[0].p, a.rk[1].p},
          {a.rv[0].p, a.rv[1].p}, a.cb};
  hipLaunchKernelGGL(k_ip_emit, grid_of(np), dim3(RP_BLOCK), 0, st, v, (const int*)a.pin_net.p,
                     (const int4*)a.ext.p, o);
  HIPCHK(c, hipGetLastError());
  for (int ax = 0; ax < 2; ax++) {
    const int m = mt[ax], r = mr[ax];
    // stable sorts: equal keys keep triplet order
    if (m > 0) {
      size_t tb = a.temp.n;
      HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(a.temp.p, tb, a.tk[ax].p, a.tk2[ax].p, a.tv[ax].p,
                                                   a.tv2[ax].p, m, 0, 2 * a.cb, st));
      hipLaunchKernelGGL(k_ip_heads<unsigned long long>, grid_of(m), dim3(RP_BLOCK), 0, st,
                         (const unsigned long long*)a.tk2[ax].p, m, a.flag.p);
      tb = a.temp.n;
      HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(a.temp.p, tb, a.flag.p, a.seg.p, m, st));
      hipLaunchKernelGGL(k_ip_segsum, grid_of(m), dim3(RP_BLOCK), 0, st, (const unsigned long long*)a.tk2[ax].p,
                         (const float*)a.tv2[ax].p, m, (const int*)a.flag.p, (const int*)a.seg.p, a.cb,
                         a.col[ax].p, a.val[ax].p, a.srow.p);
    }
    hipLaunchKernelGGL(k_ip_rowp, grid_of((long)n + 1), dim3(RP_BLOCK), 0, st, (const int*)a.srow.p,
                       (const int*)a.flag.p, (const int*)a.seg.p, m, n, a.rowp[ax].p);
    hipLaunchKernelGGL(k_ip_zero, grid_of(n), dim3(RP_BLOCK), 0, st, a.b[ax].p, n);
    if (r > 0) {
      size_t tb = a.temp.n;
      HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(a.temp.p, tb, a.rk[ax].p, a.rk2[ax].p, a.rv[ax].p,
                                                   a.rv2[ax].p, r, 0, a.cb, st));
      hipLaunchKernelGGL(k_ip_rhs, grid_of(r), dim3(RP_BLOCK), 0, st, (const unsigned*)a.rk2[ax].p,
                         (const float*)a.rv2[ax].p, r, a.b[ax].p);
    }
    HIPCHK(c, hipGetLastError());
  }
  return 0;
}

// ===========================================================================
// BiCGSTAB with the scalar recurrence on the device
// ===========================================================================
// Identity preconditioner, the published algorithm (van der Vorst 1992) in
// the variant that restarts when rho collapses (the reference calls Eigen's
// BiCGSTAB, src/initialPlace.cpp:78-86).  Both axes run in the same launches
// (blockIdx.y = axis).  FP64 dot products are reduced in a fixed order:
// per-block partials of a fixed grid, summed sequentially per thread and
// across the block by one fixed tree, so every block that needs a scalar
// computes the same value and results do not depend on timing.
namespace {

constexpr int BC_MAX_PARTS = 1024;
constexpr int BC_SLOTS = 8;  // partial-sum slots per axis
constexpr int GS_WIN = 1024;  // rows of a sorting window of the sliced-ELL copy (one workgroup)
constexpr int GS_SMAX = GS_WIN + 8;  // slices of a window at most (64 / g rows each, 7 classes of g)
enum { BC_RR0 = 0, BC_BB, BC_R0V, BC_TT, BC_TS, BC_RR, BC_R0R, BC_RST };

struct BcState {
  double rho, rho_old, alpha, w, rr, r0r, r0_sq, rhs_sq, tol2, eps2;
  int i, restarts, done, restart, zero, max_iter, started, pending;  // pending: an iteration awaits its head
  float err;
};

// One axis of a grid-wide solve.  The solve runs on the row-sorted system
// P A P^T (P x) = P b (rows sorted by length, longest first, within windows of
// GS_WIN consecutive rows, so the locality order survives at window scale):
// b, x and the work vectors are indexed by sorted position q; bo and xo are
// the right-hand side and solution in row order.  One workgroup of GS_WIN
// threads serves one window.  The matrix is a sliced ELL (after SELL-C-sigma,
// Kreutzer et al., SIAM J. Sci. Comput. 36(5), 2014) whose slices give a row
// g lanes by its length (g = 1 up to 16 entries, 2 up to 32, ... 64 above
// 512), so a slice holds 64 / g rows and no lane walks more than 16 slots
// (rows above 1024 entries excepted): lane k of a row sums entries k, k + g,
// ... in order in FP64, then the g lanes add by a fixed xor tree.  Slot m of
// lane l of a slice sits at soff + (m / 4) 256 + 4 l + m % 4, so a wave reads
// 64 x 16 contiguous bytes of values and of column positions per four slots.
struct BcAxis {
  const int* rowp;  // CSR in row order (the assembly's output)
  const int* col;
  const float* val;
  const float* bo;
  float* xo;
  const int* perm;   // sorted position -> row
  const int* slen;   // sorted position -> row length
  const int4* sinfo;  // per window, GS_SMAX slices: (first row, g | rows << 8, slots per lane, first slot)
  const int* nsl;     // per window: slices
  const float* sval;
  const int* scol;   // sorted column positions (pads: 0, never summed)
  const float* b;
  float *x, *r, *r0, *v, *p, *s, *t;
  double* parts;  // BC_SLOTS x nparts
  BcState* st;
};

struct BcArgs {
  BcAxis ax[2];
  int n, nparts;
};

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// fixed-order block sum; result valid in every thread
__device__ double bsum(double v, double* sh) {
  v = wsum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
  return r;
}

// sum of one partial slot, identical in every block
__device__ double psum(const BcAxis& a, int slot, int nparts, double* sh) {
  double v = 0.0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) v += a.parts[slot * nparts + k];
  return bsum(v, sh);
}

// y_q = sum_k val_k x_col_k for the rows of one slice (wave-uniform k of
// window win): lane l takes entries l % g, l % g + g, ... of row q0 + l / g,
// eight slots (two quads) loaded and gathered before they are summed, slots
// past the lane's count dropped by a select; then the row's g lanes add by a
// fixed xor tree.  Returns true in the lane that holds y_q (l % g == 0).
__device__ __forceinline__ bool gsell_slice(const BcAxis& a, int win, int k, const float* __restrict__ x, int& q,
                                            float& y) {
  const int4 si = a.sinfo[(size_t)win * GS_SMAX + k];
  const int q0 = __builtin_amdgcn_readfirstlane(si.x);
  const int gl = __builtin_amdgcn_readfirstlane(si.y);
  const int w = __builtin_amdgcn_readfirstlane(si.z);
  const unsigned soff = (unsigned)__builtin_amdgcn_readfirstlane(si.w);
  const int g = gl & 0xFF, rows = gl >> 8;
  const int lane = threadIdx.x & 63;
  const int rl = lane / g, kk = lane & (g - 1);
  const bool valid = rl < rows;
  q = q0 + rl;
  const int L = valid ? a.slen[q] : 0;
  const int cnt = L > kk ? (L - kk + g - 1) / g : 0;
  const float* __restrict__ pv = a.sval + soff + (unsigned)lane * 4u;
  const int* __restrict__ pc = a.scol + soff + (unsigned)lane * 4u;
  double acc = 0.0;
  if (valid) {
    int m = 0;
    for (; m + 8 <= w; m += 8) {
      const float4 v0 = *reinterpret_cast<const float4*>(pv + (unsigned)m * 64u);
      const float4 v1 = *reinterpret_cast<const float4*>(pv + (unsigned)(m + 4) * 64u);
      const int4 c0 = *reinterpret_cast<const int4*>(pc + (unsigned)m * 64u);
      const int4 c1 = *reinterpret_cast<const int4*>(pc + (unsigned)(m + 4) * 64u);
      const float vv[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
      const float xx[8] = {x[c0.x], x[c0.y], x[c0.z], x[c0.w], x[c1.x], x[c1.y], x[c1.z], x[c1.w]};
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const double t = acc + (double)vv[u] * (double)xx[u];
        acc = m + u < cnt ? t : acc;
      }
    }
    if (m < w) {  // a last quad
      const float4 v0 = *reinterpret_cast<const float4*>(pv + (unsigned)m * 64u);
      const int4 c0 = *reinterpret_cast<const int4*>(pc + (unsigned)m * 64u);
      const float vv[4] = {v0.x, v0.y, v0.z, v0.w};
      const float xx[4] = {x[c0.x], x[c0.y], x[c0.z], x[c0.w]};
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const double t = acc + (double)vv[u] * (double)xx[u];
        acc = m + u < cnt ? t : acc;
      }
    }
  }
  for (int o = g >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  y = (float)acc;
  return valid && kk == 0;
}

// Kernel bodies take (axis, block), so that one body serves a launch per
// context and a launch over many contexts (a job table); g.nparts blocks
// cover an axis either way, block b the row windows b, b + nparts, ..., so
// every dot product has one partial per block.
// f(q, y_q) for every sorted row of the block's windows, the windows' slices
// dealt to its waves in turn
template <typename F>
__device__ __forceinline__ void gsell_rows(const BcArgs& g, const BcAxis& a, int blk, const float* __restrict__ x,
                                           F&& f) {
  const int nwin = (g.n + GS_WIN - 1) / GS_WIN, wave = threadIdx.x >> 6;
  for (int win = blk; win < nwin; win += g.nparts) {
    const int ns = a.nsl[win];
    for (int k = wave; k < ns; k += GS_WIN / 64) {
      int q;
      float y;
      if (gsell_slice(a, win, k, x, q, y)) f(q, y);
    }
  }
}

// The row-wise kernels (p, s, xr, unperm) run RP_BLOCK threads a window,
// four rows a thread: ld(q, u) for the four rows first, then st(q, u), so a
// thread has its four rows' loads in flight at once.
static_assert(GS_WIN == 4 * RP_BLOCK, "four rows a thread");
template <typename Ld, typename St>
__device__ __forceinline__ void bc_rows4(const BcArgs& g, int blk, Ld&& ld, St&& st) {
  const int nwin = (g.n + GS_WIN - 1) / GS_WIN;
  for (int win = blk; win < nwin; win += g.nparts) {
    const int q0 = win * GS_WIN + (int)threadIdx.x;
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (q0 + u * RP_BLOCK < g.n) ld(q0 + u * RP_BLOCK, u);
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (q0 + u * RP_BLOCK < g.n) st(q0 + u * RP_BLOCK, u);
  }
}

// r = b - A x ; r0 = r ; v = p = 0 ; partials r.r and b.b
__device__ __forceinline__ void bc_init_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[GS_WIN / 64];
  const BcAxis& a = g.ax[axis];
  double d0 = 0.0, d1 = 0.0;
  gsell_rows(g, a, blk, a.x, [&](int i, float t) {
    float r = (float)((double)a.b[i] + -1.0 * (double)t);
    a.r[i] = r;
    a.r0[i] = r;
    a.v[i] = 0.0f;
    a.p[i] = 0.0f;
    d0 += (double)r * (double)r;
    d1 += (double)a.b[i] * (double)a.b[i];
  });
  d0 = bsum(d0, sh);
  d1 = bsum(d1, sh);
  if (threadIdx.x == 0) {
    a.parts[BC_RR0 * g.nparts + blk] = d0;
    a.parts[BC_BB * g.nparts + blk] = d1;
    a.parts[BC_RR * g.nparts + blk] = 0.0;  // xr rewrites those of the blocks with rows
    a.parts[BC_R0R * g.nparts + blk] = 0.0;
  }
}

// Scalar head of an iteration: finishes the previous iteration (or the
// set-up), decides termination, and prepares rho / the restart test.
__device__ __forceinline__ void bc_head_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[1024 / 64];
  const BcAxis& a = g.ax[axis];
  BcState& S = *a.st;
  if (S.done || (S.started && !S.pending)) return;  // idempotent: nothing to close
  double rr, r0r;
  int i;
  if (!S.started) {
    double r0_sq = psum(a, BC_RR0, g.nparts, sh);
    double rhs_sq = psum(a, BC_BB, g.nparts, sh);
    rr = r0r = r0_sq;
    i = 0;
    if (threadIdx.x == 0) {
      S.r0_sq = r0_sq;
      S.rhs_sq = rhs_sq;
      S.zero = rhs_sq == 0.0;
      S.tol2 = S.tol2 * rhs_sq;  // tol^2 was stored by the host
      S.rho = S.alpha = S.w = 1.0;
      S.restarts = 0;
      S.started = 1;
    }
  } else {
    rr = psum(a, BC_RR, g.nparts, sh);
    r0r = psum(a, BC_R0R, g.nparts, sh);
    i = S.i + 1;
  }
  if (threadIdx.x != 0) return;
  const bool zero = S.zero;
  S.pending = 0;
  S.rr = rr;
  S.r0r = r0r;
  S.i = i;
  if (zero || !(rr > S.tol2 && i < S.max_iter)) {
    S.done = 1;
    S.err = zero ? 0.0f : (float)sqrt(rr / S.rhs_sq);
    return;
  }
  S.rho_old = S.rho;
  S.rho = r0r;
  S.restart = fabs(S.rho) < S.eps2 * S.r0_sq;
}

// restart: r = b - A x ; r0 = r ; partial r.r
__device__ __forceinline__ void bc_restart_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[GS_WIN / 64];
  const BcAxis& a = g.ax[axis];
  const BcState& S = *a.st;
  if (S.done || !S.restart) return;
  double d0 = 0.0;
  gsell_rows(g, a, blk, a.x, [&](int i, float t) {
    float r = (float)((double)a.b[i] + -1.0 * (double)t);
    a.r[i] = r;
    a.r0[i] = r;
    d0 += (double)r * (double)r;
  });
  d0 = bsum(d0, sh);
  if (threadIdx.x == 0) a.parts[BC_RST * g.nparts + blk] = d0;
}

// p = r + beta (p - w v)   (after a restart, rho = r0_sq = r.r)
__device__ __forceinline__ void bc_p_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[GS_WIN / 64];
  const BcAxis& a = g.ax[axis];
  BcState& S = *a.st;
  if (S.done) return;
  double rho = S.rho;
  if (S.restart) {
    rho = psum(a, BC_RST, g.nparts, sh);
    __syncthreads();  // every thread has read S before block 0 writes it
    if (blk == 0 && threadIdx.x == 0) {  // no other block reads these here
      S.rho = S.r0_sq = rho;
      if (S.restarts++ == 0) S.i = 0;
    }
  }
  const double beta = (rho / S.rho_old) * (S.alpha / S.w), w = S.w;
  float rv[4], pv[4], vv[4];
  bc_rows4(
      g, blk,
      [&](int i, int u) {
        rv[u] = a.r[i];
        pv[u] = a.p[i];
        vv[u] = a.v[i];
      },
      [&](int i, int u) { a.p[i] = (float)((double)rv[u] + beta * ((double)pv[u] - w * (double)vv[u])); });
}

// v = A p ; partial r0.v
__device__ __forceinline__ void bc_v_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[GS_WIN / 64];
  const BcAxis& a = g.ax[axis];
  if (a.st->done) return;
  double d = 0.0;
  gsell_rows(g, a, blk, a.p, [&](int i, float v) {
    a.v[i] = v;
    d += (double)a.r0[i] * (double)v;
  });
  d = bsum(d, sh);
  if (threadIdx.x == 0) a.parts[BC_R0V * g.nparts + blk] = d;
}

// alpha = rho / (r0.v) ; s = r - alpha v
__device__ __forceinline__ void bc_s_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[GS_WIN / 64];
  const BcAxis& a = g.ax[axis];
  BcState& S = *a.st;
  if (S.done) return;
  const double alpha = S.rho / psum(a, BC_R0V, g.nparts, sh);
  __syncthreads();
  if (blk == 0 && threadIdx.x == 0) S.alpha = alpha;
  float rv[4], vv[4];
  bc_rows4(
      g, blk,
      [&](int i, int u) {
        rv[u] = a.r[i];
        vv[u] = a.v[i];
      },
      [&](int i, int u) { a.s[i] = (float)((double)rv[u] + -alpha * (double)vv[u]); });
}

// t = A s ; partials t.t and t.s
__device__ __forceinline__ void bc_t_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[GS_WIN / 64];
  const BcAxis& a = g.ax[axis];
  if (a.st->done) return;
  double tt = 0.0, ts = 0.0;
  gsell_rows(g, a, blk, a.s, [&](int i, float t) {
    a.t[i] = t;
    tt += (double)t * (double)t;
    ts += (double)t * (double)a.s[i];
  });
  tt = bsum(tt, sh);
  ts = bsum(ts, sh);
  if (threadIdx.x == 0) {
    a.parts[BC_TT * g.nparts + blk] = tt;
    a.parts[BC_TS * g.nparts + blk] = ts;
  }
}

// w = (t.s)/(t.t) ; x += alpha p + w s ; r = s - w t ; partials r.r and r0.r
__device__ __forceinline__ void bc_xr_body(const BcArgs& g, int axis, int blk) {
  __shared__ double sh[GS_WIN / 64];
  const BcAxis& a = g.ax[axis];
  BcState& S = *a.st;
  if (S.done) return;
  const double tt = psum(a, BC_TT, g.nparts, sh);
  const double ts = psum(a, BC_TS, g.nparts, sh);
  const double w = tt > 0.0 ? ts / tt : 0.0, alpha = S.alpha;
  __syncthreads();
  if (blk == 0 && threadIdx.x == 0) {
    S.w = w;
    S.pending = 1;
  }
  double d0 = 0.0, d1 = 0.0;
  float xv[4], pv[4], sv[4], tv[4], r0v[4];
  bc_rows4(
      g, blk,
      [&](int i, int u) {
        xv[u] = a.x[i];
        pv[u] = a.p[i];
        sv[u] = a.s[i];
        tv[u] = a.t[i];
        r0v[u] = a.r0[i];
      },
      [&](int i, int u) {
        a.x[i] = (float)((double)xv[u] + alpha * (double)pv[u] + w * (double)sv[u]);
        float r = (float)((double)sv[u] - w * (double)tv[u]);
        a.r[i] = r;
        d0 += (double)r * (double)r;
        d1 += (double)r0v[u] * (double)r;
      });
  d0 = bsum(d0, sh);
  d1 = bsum(d1, sh);
  if (threadIdx.x == 0) {
    a.parts[BC_RR * g.nparts + blk] = d0;
    a.parts[BC_R0R * g.nparts + blk] = d1;
  }
}

// the solution back in row order (a zero right-hand side gives x = 0)
__device__ __forceinline__ void bc_unperm_body(const BcArgs& g, int axis, int blk) {
  const BcAxis& a = g.ax[axis];
  const bool zero = a.st->zero != 0;
  float xv[4];
  int pm[4];
  bc_rows4(
      g, blk,
      [&](int i, int u) {
        xv[u] = a.x[i];
        pm[u] = a.perm[i];
      },
      [&](int i, int u) { a.xo[pm[u]] = zero ? 0.0f : xv[u]; });
}

// one launch per context (blockIdx.y = axis) and one over many contexts
// (entry blockIdx.x of tab = (context, block))
#define BC_KERNELS(NAME, NT)                                                             \
  __global__ void __launch_bounds__(NT) k_bc_##NAME(BcArgs g) {                         \
    bc_##NAME##_body(g, blockIdx.y, blockIdx.x);                                         \
  }                                                                                      \
  __global__ void __launch_bounds__(NT) kb_bc_##NAME(const BcArgs* gs, const int2* tab) { \
    const int2 e = tab[blockIdx.x];                                                      \
    bc_##NAME##_body(gs[e.x], blockIdx.y, e.y);                                          \
  }
BC_KERNELS(init, GS_WIN)
BC_KERNELS(head, 1024)
BC_KERNELS(restart, GS_WIN)
BC_KERNELS(p, RP_BLOCK)
BC_KERNELS(v, GS_WIN)
BC_KERNELS(s, RP_BLOCK)
BC_KERNELS(t, GS_WIN)
BC_KERNELS(xr, RP_BLOCK)
BC_KERNELS(unperm, RP_BLOCK)

// lanes per row of a row of L entries: no lane walks more than 16 slots
__device__ __forceinline__ int gsell_lanes(int L) {
  int g = 1;
  while (g < 64 && L > 16 * g) g <<= 1;
  return g;
}

// Sliced-ELL copy of one axis' CSR for the grid-wide solve, part 1, one
// workgroup per window of GS_WIN rows: the window's rows sorted by length,
// longest first, ties in row order (a bitonic sort of length-major keys that
// hold the row's place in the window, so the order is unique; lengths above
// 65534 count as equal), the sorted position -> row map and its inverse, the
// sorted lengths, and the window's slices: consecutive sorted rows of one
// lane class g, 64 / g of them (fewer at a class's end), each as wide as its
// longest row needs (slots per lane rounded up to a multiple of 4).  Slice
// slot offsets are relative to the window here; part 2 adds the window's base.
__global__ void __launch_bounds__(GS_WIN) k_gsell_order(const int* __restrict__ rowp, int n, int* __restrict__ perm,
                                                       int* __restrict__ iperm, int* __restrict__ slen,
                                                       int* __restrict__ qslice, int4* __restrict__ sinfo,
                                                       int* __restrict__ nsl, int* __restrict__ wtot) {
  __shared__ unsigned key[GS_WIN];
  __shared__ int len_s[GS_WIN];
  const int tid = threadIdx.x, w0 = blockIdx.x * GS_WIN;
  const int nw = min(GS_WIN, n - w0);
  const int i = w0 + tid;
  const int len = i < n ? rowp[i + 1] - rowp[i] : 0;
  key[tid] = ((i < n ? 0xFFFFu - (unsigned)min(len, 0xFFFF) : 0x10000u) << 10) | (unsigned)tid;
  __syncthreads();
  for (int k = 2; k <= GS_WIN; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int o = tid ^ j;
      if (o > tid) {
        const unsigned x = key[tid], y = key[o];
        if (((tid & k) == 0) == (x > y)) {
          key[tid] = y;
          key[o] = x;
        }
      }
      __syncthreads();
    }
  const int q = w0 + tid;
  const int src = w0 + (int)(key[tid] & (GS_WIN - 1));
  int l = 0;
  if (tid < nw) {  // rows past n sort last
    l = rowp[src + 1] - rowp[src];
    perm[q] = src;
    iperm[src] = q;
    slen[q] = l;
  }
  // slices: within each lane class (contiguous, larger g first, as the rows
  // are sorted by length) runs of 64 / g rows from the class's first row
  __shared__ int cstart[8], sbase[8], scan[GS_WIN];
  if (tid < 8) cstart[tid] = nw;
  __syncthreads();
  const int cg = tid < nw ? 31 - __clz(gsell_lanes(l)) : 0;  // log2 g
  len_s[tid] = cg;
  __syncthreads();
  if (tid < nw && (tid == 0 || len_s[tid - 1] != cg)) cstart[cg] = tid;
  __syncthreads();
  if (tid == 0) {  // first slice of each class, classes in position order (g = 64 first)
    int k = 0;
    for (int c = 6; c >= 0; c--) {
      sbase[c] = k;
      if (cstart[c] < nw) {
        int e = nw;  // the class ends where the next present class starts
        for (int d = c - 1; d >= 0; d--)
          if (cstart[d] < nw) {
            e = cstart[d];
            break;
          }
        const int h = 64 >> c;
        k += (e - cstart[c] + h - 1) / h;
      }
    }
    sbase[7] = k;  // slices of the window
  }
  __syncthreads();
  int kslice = 0, slots = 0, rows = 0, w = 0;
  bool head = false;
  if (tid < nw) {
    const int g = 1 << cg, h = 64 >> cg, loc = tid - cstart[cg];
    int e = nw;
    for (int d = cg - 1; d >= 0; d--)
      if (cstart[d] < nw) {
        e = cstart[d];
        break;
      }
    kslice = sbase[cg] + loc / h;
    head = loc % h == 0;
    if (head) {  // the slice's first row is its longest
      rows = min(h, e - tid);
      w = ((l + g - 1) / g + 3) & ~3;
      slots = 64 * w;
    }
  }
  // exclusive scan of the slices' slots in position order (Hillis-Steele)
  scan[tid] = slots;
  __syncthreads();
  for (int o = 1; o < GS_WIN; o <<= 1) {
    const int v = tid >= o ? scan[tid - o] : 0;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  if (head)
    sinfo[(size_t)blockIdx.x * GS_SMAX + kslice] = make_int4(q, (1 << cg) | (rows << 8), w, scan[tid] - slots);
  if (tid < nw) qslice[q] = kslice;
  if (tid == 0) {
    nsl[blockIdx.x] = sbase[7];
    wtot[blockIdx.x] = scan[GS_WIN - 1];
  }
}

// Part 2, one workgroup per window: the window's slot base (the windows
// before it) added to its slices' offsets, the entries with sorted column
// positions (a lane's slots past its count: value 0, column 0), and b and x in
// sorted order.
__global__ void __launch_bounds__(GS_WIN) k_gsell_fill(const int* __restrict__ rowp, const int* __restrict__ col,
                                                      const float* __restrict__ val, const float* __restrict__ b,
                                                      const float* __restrict__ x, const int* __restrict__ perm,
                                                      const int* __restrict__ iperm, const int* __restrict__ slen,
                                                      const int* __restrict__ qslice, const int* __restrict__ nsl,
                                                      const int* __restrict__ wtot, int n, int4* __restrict__ sinfo,
                                                      float* __restrict__ sval, int* __restrict__ scol,
                                                      float* __restrict__ bs, float* __restrict__ xs) {
  __shared__ int base;
  const int tid = threadIdx.x, w0 = blockIdx.x * GS_WIN;
  const int nw = min(GS_WIN, n - w0);
  if (tid == 0) base = 0;
  __syncthreads();
  int part = 0;
  for (int k = tid; k < (int)blockIdx.x; k += GS_WIN) part += wtot[k];
  if (part) atomicAdd(&base, part);  // integer sum: any order gives the same total
  __syncthreads();
  int4* si = sinfo + (size_t)blockIdx.x * GS_SMAX;
  const int ns = nsl[blockIdx.x];
  int4 mine = make_int4(0, 0, 0, 0);
  if (tid < nw) mine = si[qslice[w0 + tid]];
  __syncthreads();  // every thread has read its slice before they are rebased
  for (int k = tid; k < ns; k += GS_WIN) si[k].w += base;
  if (tid >= nw) return;
  const int q = w0 + tid;
  const int g = mine.y & 0xFF, w = mine.z;
  const int rl = q - mine.x;
  const int soff = base + mine.w;
  const int len = slen[q];
  const int rb = rowp[perm[q]];
  for (int kk = 0; kk < g; kk++) {
    const int lane = rl * g + kk;
    for (int m = 0; m < w; m += 4) {  // one quad: loads, then gathers, then 16-byte stores
      float v4[4];
      int c4[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = (m + u) * g + kk;  // entry of the row in this slot
        v4[u] = j < len ? val[rb + j] : 0.0f;
        c4[u] = j < len ? col[rb + j] : -1;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) c4[u] = c4[u] >= 0 ? iperm[c4[u]] : 0;
      const int e = soff + (m >> 2) * 256 + lane * 4;
      *reinterpret_cast<float4*>(sval + e) = make_float4(v4[0], v4[1], v4[2], v4[3]);
      *reinterpret_cast<int4*>(scol + e) = make_int4(c4[0], c4[1], c4[2], c4[3]);
    }
  }
  bs[q] = b[perm[q]];
  xs[q] = x[perm[q]];
}

// done flags of ng grid runs (both axes) into one array, read back in one copy
__global__ void __launch_bounds__(RP_BLOCK) k_bc_done_gather(const BcArgs* gs, int ng, int* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 2 * ng) out[k] = gs[k >> 1].ax[k & 1].st->done;
}

// ---------------------------------------------------------------------------
// Small systems: the whole solve in one workgroup per axis
// ---------------------------------------------------------------------------
// For n <= BCS_THREADS * BCS_MAX_R rows the recurrence above (same formulas,
// same restart rule, same termination test) runs in one 1024-thread
// workgroup per axis, with no launch or host poll per iteration.
//
// The system is solved in a row order sorted by length (a stable sort,
// longest first; the permuted system P A P^T (P x) = P b has the same
// iterates).  Thread tid owns sorted rows q = tid + 1024 k (k < R): r, v and
// t of its rows live in registers; p and s (gathered by the products), r0
// and x live in LDS.  The matrix is held as sliced ELL (SELL-C-sigma,
// Kreutzer et al., SIAM J. Sci. Comput. 36(5), 2014) with C = 64: each
// 64-row slice, one per wave, stores its rows slot-major at the slice's
// width (a multiple of 8 slots), so every load instruction of a wave reads 64
// consecutive entries, and the sort keeps rows of one slice at nearly the same
// length; within a slice each lane's slots are grouped in fours (16-byte value
// loads, 8-byte loads of 16-bit column positions).  The
// entries of a row keep their CSR order and are summed in FP64 in that order;
// dot products are per-thread FP64 sums over its rows in k order, then one
// fixed tree (wave xor, then the waves in order), so every thread holds the
// same scalars and takes the same branches.
constexpr int BCS_THREADS = 1024;
constexpr int BCS_MAX_R = 8;
constexpr int BCS_PASSES = 16;  // iterations per launch
constexpr int SELL_C = 64;
constexpr int SELL_PAD = 8;  // slice widths are multiples of this many slots
static_assert(BCS_THREADS * BCS_MAX_R <= 65536, "column positions are stored in 16 bits");

struct SellAxis {
  const int* rowp;
  const int* col;
  const float* val;
  const float* b;
  float* x;
  const int* perm;    // sorted position -> row
  const int* slen;    // sorted position -> row length
  const int* width;   // per slice: entries, SELL_C x its slots (its first row's length, rounded up)
  const int* soff;    // per slice: first entry
  const float* sval;  // SELL entries, slot-major per slice
  const uint16_t* scol;  // sorted column positions, 16 bits (pads: the row itself)
  BcState* st;
  float* keep;        // 5 n: p, r0, x, r, v carried from one launch to the next
  struct BcsCarry* carry;
};

// Scalars carried from one launch of a solve to the next.
struct BcsCarry {
  double rho, alpha, w, r0_sq, rr, r0r, rhs_sq;
  int it, restarts, zero;
};

struct SellArgs {
  SellAxis ax[2];
  int n;
};

// prep: descending-length sort keys
__global__ void __launch_bounds__(RP_BLOCK) k_sell_keys(const int* __restrict__ rowp, int n,
                                                        unsigned* __restrict__ key, int* __restrict__ id) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = 0xFFFFu - (unsigned)min(rowp[i + 1] - rowp[i], 0xFFFF);
  id[i] = i;
}

// prep: inverse order, sorted lengths, slice widths (rows past n are empty;
// widths rounded up to a multiple of SELL_PAD slots)
__global__ void __launch_bounds__(RP_BLOCK) k_sell_meta(const int* __restrict__ rowp, const int* __restrict__ perm,
                                                        int n, int nslices, int* __restrict__ iperm,
                                                        int* __restrict__ slen, int* __restrict__ wslots) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) {
    const int i = perm[q];
    iperm[i] = q;
    slen[q] = rowp[i + 1] - rowp[i];
  }
  if (q < nslices) {
    const int q0 = q * SELL_C;
    const int w = q0 < n ? rowp[perm[q0] + 1] - rowp[perm[q0]] : 0;
    wslots[q] = SELL_C * ((w + SELL_PAD - 1) / SELL_PAD * SELL_PAD);
  }
}

// prep: SELL entries; a slice's slots past a row's length hold val 0 and the
// row's own column (never summed)
__global__ void __launch_bounds__(RP_BLOCK) k_sell_fill(const int* __restrict__ rowp, const int* __restrict__ col,
                                                        const float* __restrict__ val, const int* __restrict__ perm,
                                                        const int* __restrict__ iperm, const int* __restrict__ soff,
                                                        int n, int nrows_pad, float* __restrict__ sval,
                                                        uint16_t* __restrict__ scol) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrows_pad) return;
  const int s = q / SELL_C, lane = q % SELL_C;
  const int q0 = s * SELL_C;
  const int w0 = q0 < n ? rowp[perm[q0] + 1] - rowp[perm[q0]] : 0;
  const int w = (w0 + SELL_PAD - 1) / SELL_PAD * SELL_PAD;
  const int rb = q < n ? rowp[perm[q]] : 0;
  const int len = q < n ? rowp[perm[q] + 1] - rb : 0;
  const int own = q < n ? q : 0;
  for (int j = 0; j < w; j++) {
    const int e = soff[s] + (j / 4) * (4 * SELL_C) + lane * 4 + (j % 4);
    const bool in = j < len;
    sval[e] = in ? val[rb + j] : 0.0f;
    scol[e] = (uint16_t)(in ? iperm[col[rb + j]] : own);
  }
}

// DPP move of a double (both halves), CTRL a dpp_ctrl encoding
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Wave sum on the DPP network: lane pairs (xor 1, xor 2), then the mirrored
// half-row and row, which gives every lane of a 16-lane row the same row sum
// (each step adds two values that are the same pair in every lane, and a
// floating-point add is commutative); the four row sums are then added in
// row order.  The result is identical in every lane.
__device__ __forceinline__ double wsum_dpp(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  return ((readlane_f64(v, 0) + readlane_f64(v, 16)) + readlane_f64(v, 32)) + readlane_f64(v, 48);
}

// sums of a and b over the workgroup, identical in every thread
__device__ __forceinline__ void bcs_sum2(double& a, double& b, double* red) {
  a = wsum_dpp(a);
  b = wsum_dpp(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  double x = 0.0, y = 0.0;
  for (int k = 0; k < BCS_THREADS / 64; k++) {
    x += red[2 * k];
    y += red[2 * k + 1];
  }
  a = x;
  b = y;
}

// y = A X over the thread's rows (X in LDS, sorted order).  A round loads
// SB slots (SB divides SELL_PAD, so slice widths are whole rounds) and issues
// its loads and gathers unconditionally; an entry past the row's length is
// dropped by a select, so each row's sum is exactly the CSR-order sum of its
// entries.
template <int R, int SB = 8>
__device__ __forceinline__ void bcs_spmv(const SellAxis& a, const int (&len)[R], const float* X, float (&y)[R]) {
  static_assert(SELL_PAD % SB == 0 && SB % 4 == 0, "rounds must tile the slice widths in quads");
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < R; k++) {
    const unsigned s = wave + (BCS_THREADS / SELL_C) * k;
    const int w = __builtin_amdgcn_readfirstlane(a.width[s]) / SELL_C;
    const unsigned base = __builtin_amdgcn_readfirstlane(a.soff[s]);
    const float* __restrict__ pv = a.sval + base;
    const uint16_t* __restrict__ pc = a.scol + base;
    double acc = 0.0;
    for (int j = 0; j < w; j += SB) {
      float vv[SB], xx[SB];
      int cc[SB];
#pragma unroll
      for (int u = 0; u < SB; u += 4) {
        const unsigned e = (unsigned)(j + u) * SELL_C + lane * 4;
        const float4 v4 = *reinterpret_cast<const float4*>(pv + e);
        const ushort4 c4 = *reinterpret_cast<const ushort4*>(pc + e);
        vv[u] = v4.x;
        vv[u + 1] = v4.y;
        vv[u + 2] = v4.z;
        vv[u + 3] = v4.w;
        cc[u] = c4.x;
        cc[u + 1] = c4.y;
        cc[u + 2] = c4.z;
        cc[u + 3] = c4.w;
      }
#pragma unroll
      for (int u = 0; u < SB; u++) xx[u] = X[cc[u]];
#pragma unroll
      for (int u = 0; u < SB; u++) {
        const double t = acc + (double)vv[u] * (double)xx[u];
        acc = j + u < len[k] ? t : acc;
      }
    }
    y[k] = (float)acc;
  }
}

// Passes [first, first + npass) of the solve; the first launch (first = 0)
// starts it, later ones resume from the carried state.  Launches stay short
// (a queue shared with other work is not held for a whole solve).
template <int R>
__device__ __forceinline__ void bcs_body(const SellArgs& g, int axis, int first, int npass) {
  extern __shared__ float bcs_lds[];
  __shared__ double red[2 * (BCS_THREADS / 64)];
  const SellAxis& a = g.ax[axis];
  BcState& S = *a.st;
  if (S.done) return;
  const int n = g.n, tid = threadIdx.x;
  float* P = bcs_lds;           // p
  float* Q = bcs_lds + n;       // x for r = b - A x, otherwise s
  float* R0 = bcs_lds + 2 * n;  // r0
  float* X = bcs_lds + 3 * n;   // x
  int len[R];
  float r[R], v[R], t[R];
  float* keep_p = a.keep;
  float* keep_r0 = a.keep + n;
  float* keep_x = a.keep + 2 * n;
  float* keep_r = a.keep + 3 * n;
  float* keep_v = a.keep + 4 * n;
  double d0 = 0.0, d1 = 0.0;
  double rhs_sq, r0_sq, rr, r0r, rho, alpha, w;
  int it, restarts;
  bool zero;
  if (first == 0) {
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      len[k] = q < n ? a.slen[q] : 0;
      if (q < n) Q[q] = X[q] = a.x[a.perm[q]];
    }
    __syncthreads();
    // r = b - A x ; r0 = r ; v = p = 0
    bcs_spmv<R>(a, len, Q, t);
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      r[k] = v[k] = 0.0f;
      if (q < n) {
        const float bi = a.b[a.perm[q]];
        const float ri = (float)((double)bi + -1.0 * (double)t[k]);
        r[k] = R0[q] = ri;
        P[q] = 0.0f;
        d0 += (double)ri * (double)ri;
        d1 += (double)bi * (double)bi;
      }
    }
    bcs_sum2(d0, d1, red);
    rhs_sq = d1;
    zero = rhs_sq == 0.0;
    r0_sq = rr = r0r = d0;
    rho = alpha = w = 1.0;
    it = restarts = 0;
  } else {
    const BcsCarry& C = *a.carry;
    rho = C.rho;
    alpha = C.alpha;
    w = C.w;
    r0_sq = C.r0_sq;
    rr = C.rr;
    r0r = C.r0r;
    rhs_sq = C.rhs_sq;
    it = C.it;
    restarts = C.restarts;
    zero = C.zero != 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      len[k] = q < n ? a.slen[q] : 0;
      r[k] = v[k] = 0.0f;
      if (q < n) {
        P[q] = keep_p[q];
        R0[q] = keep_r0[q];
        X[q] = keep_x[q];
        r[k] = keep_r[q];
        v[k] = keep_v[q];
      }
    }
    __syncthreads();
  }
  const double tol2 = S.tol2 * rhs_sq, eps2 = S.eps2;  // S.tol2 holds tol^2
  const int max_iter = S.max_iter;
  int done = 0;
  // i restarts from 0 at most once, so 2 max_iter + 1 iterations bound the solve
  const int guard = 2 * max(0, max_iter) + 2;
  const int last = min(first + npass, guard);
  int pass = first;
  for (; pass < last; pass++) {
    if (zero || !(rr > tol2 && it < max_iter)) {
      done = 1;
      break;
    }
    const double rho_old = rho;
    rho = r0r;
    if (fabs(rho) < eps2 * r0_sq) {  // restart: r = b - A x ; r0 = r ; rho = r.r
#pragma unroll
      for (int k = 0; k < R; k++) {
        const int q = tid + BCS_THREADS * k;
        if (q < n) Q[q] = X[q];
      }
      __syncthreads();
      bcs_spmv<R>(a, len, Q, t);
      d0 = d1 = 0.0;
#pragma unroll
      for (int k = 0; k < R; k++) {
        const int q = tid + BCS_THREADS * k;
        if (q < n) {
          const float ri = (float)((double)a.b[a.perm[q]] + -1.0 * (double)t[k]);
          r[k] = R0[q] = ri;
          d0 += (double)ri * (double)ri;
        }
      }
      bcs_sum2(d0, d1, red);
      rho = r0_sq = d0;
      if (restarts++ == 0) it = 0;
    }
    const double beta = (rho / rho_old) * (alpha / w);
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      if (q < n) P[q] = (float)((double)r[k] + beta * ((double)P[q] - w * (double)v[k]));
    }
    __syncthreads();
    // v = A p ; alpha = rho / r0.v
    bcs_spmv<R>(a, len, P, v);
    d0 = d1 = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      if (q < n) d0 += (double)R0[q] * (double)v[k];
    }
    bcs_sum2(d0, d1, red);
    alpha = rho / d0;
    // s = r - alpha v ; t = A s ; w = t.s / t.t
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      if (q < n) Q[q] = (float)((double)r[k] + -alpha * (double)v[k]);
    }
    __syncthreads();
    bcs_spmv<R>(a, len, Q, t);
    double tt = 0.0, ts = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      if (q < n) {
        tt += (double)t[k] * (double)t[k];
        ts += (double)t[k] * (double)Q[q];
      }
    }
    bcs_sum2(tt, ts, red);
    w = tt > 0.0 ? ts / tt : 0.0;
    // x += alpha p + w s ; r = s - w t
    d0 = d1 = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      if (q < n) {
        X[q] = (float)((double)X[q] + alpha * (double)P[q] + w * (double)Q[q]);
        const float ri = (float)((double)Q[q] - w * (double)t[k]);
        r[k] = ri;
        d0 += (double)ri * (double)ri;
        d1 += (double)R0[q] * (double)ri;
      }
    }
    bcs_sum2(d0, d1, red);
    rr = d0;
    r0r = d1;
    it++;
  }
  if (!done) {
    if (pass >= guard) return;  // not finished within the bound: the host reports it
    // carry the state to the next launch
#pragma unroll
    for (int k = 0; k < R; k++) {
      const int q = tid + BCS_THREADS * k;
      if (q < n) {
        keep_p[q] = P[q];
        keep_r0[q] = R0[q];
        keep_x[q] = X[q];
        keep_r[q] = r[k];
        keep_v[q] = v[k];
      }
    }
    if (tid == 0) {
      BcsCarry& C = *a.carry;
      C.rho = rho;
      C.alpha = alpha;
      C.w = w;
      C.r0_sq = r0_sq;
      C.rr = rr;
      C.r0r = r0r;
      C.rhs_sq = rhs_sq;
      C.it = it;
      C.restarts = restarts;
      C.zero = zero;
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int q = tid + BCS_THREADS * k;
    if (q < n) a.x[a.perm[q]] = zero ? 0.0f : X[q];
  }
  if (tid == 0) {
    S.done = done;
    S.zero = zero;
    S.i = it;
    S.restarts = restarts;
    S.pending = pass;  // iterations run (diagnostics)
    S.rr = rr;
    S.rhs_sq = rhs_sq;
    S.err = zero ? 0.0f : (float)sqrt(rr / rhs_sq);
  }
}

// one context per launch (blockIdx.x = axis) and many (blockIdx.x = 2 context + axis)
template <int R>
__global__ void __launch_bounds__(BCS_THREADS) k_bc_small(SellArgs g, int first, int npass) {
  bcs_body<R>(g, blockIdx.x, first, npass);
}
template <int R>
__global__ void __launch_bounds__(BCS_THREADS) kb_bc_small(const SellArgs* gs, int first, int npass) {
  bcs_body<R>(gs[blockIdx.x >> 1], blockIdx.x & 1, first, npass);
}

// rows per thread of the one-workgroup solve, 0 when n is too large for it
inline int bcs_rows_per_thread(int n) {
  for (int R = 1; R <= BCS_MAX_R; R *= 2)
    if (n <= BCS_THREADS * R) return R;
  return 0;
}

// Work arrays of the sorted SELL copy of one axis.
struct SellBufs {
  DevBuf<unsigned> key, key2;
  DevBuf<int> id, perm, iperm, slen, wslots, soff;
  DevBuf<float> sval;
  DevBuf<uint16_t> scol;
  DevBuf<char> temp;
  DevBuf<int> scol32, wtot, nsl, qslice;  // grid-wide solve: 32-bit columns, slots and slices per window, slice of a row
  DevBuf<int4> sinfo;                     // grid-wide solve: slices of each window
};

// ---------------------------------------------------------------------------
// InitialPlace iteration kernels: warm start, instance-box HPWL, updateCoordi
// ---------------------------------------------------------------------------
struct IpMove {
  int n, n_net, n_pin;
  int* icx;
  int* icy;
  int* dblx;  // DB lower-left of movable k (what PlacerBase::hpwl reads)
  int* dbly;
  const int* hx;  // half density-size used by Instance::setCenterLocation
  const int* hy;
  const int* mw;
  const int* mh;
  const int* net_start;
  const int* pin_k;
  const int* offx;
  const int* offy;
  int* pin_cx;
  int* pin_cy;
  const int4* sbox;  // per pin: box of a fixed ITerm / placed BTerm, or (MAX, MAX, MIN, MIN)
  long long* hparts;
  float* x0;
  float* x1;
};

__global__ void __launch_bounds__(RP_BLOCK) k_ip_warm(IpMove m) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m.n) return;
  m.x0[k] = (float)m.icx[k];
  m.x1[k] = (float)m.icy[k];
}

// PlacerBase::hpwl over instance boxes (src/placerBase.cpp:868-876, 480-500)
__global__ void __launch_bounds__(RP_BLOCK) k_ip_hpwl(IpMove m) {
  __shared__ long long sh[RP_BLOCK / 64];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  long long h = 0;
  if (e < m.n_net) {
    int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
    for (int q = m.net_start[e]; q < m.net_start[e + 1]; q++) {
      int k = m.pin_k[q];
      int4 b;
      if (k >= 0) {
        b.x = m.dblx[k];
        b.y = m.dbly[k];
        b.z = b.x + m.mw[k];
        b.w = b.y + m.mh[k];
      } else {
        b = m.sbox[q];
      }
      lx = min(lx, b.x);
      ly = min(ly, b.y);
      ux = max(ux, b.z);
      uy = max(uy, b.w);
    }
    h = (long long)((int)((unsigned)ux - (unsigned)lx) + (int)((unsigned)uy - (unsigned)ly));
  }
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int w = 0; w < RP_BLOCK / 64; w++) t += sh[w];
    m.hparts[blockIdx.x] = t;
  }
}

__global__ void k_ip_hpwl_final(const long long* parts, int np, long long* out) {
  __shared__ long long sh[1024 / 64];
  long long h = 0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) h += parts[k];
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w];
    *out = t;
  }
}

// updateCoordi (src/initialPlace.cpp:380-385): centres from the solution,
// Instance::setCenterLocation and the DB location
__global__ void __launch_bounds__(RP_BLOCK) k_ip_update(IpMove m) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m.n) return;
  int x = (int)m.x0[k], y = (int)m.x1[k];
  m.icx[k] = x;
  m.icy[k] = y;
  m.dblx[k] = x - m.hx[k];
  m.dbly[k] = y - m.hy[k];
}

// Pin::updateLocation for pins of movable instances
__global__ void __launch_bounds__(RP_BLOCK) k_ip_pins(IpMove m) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m.n_pin) return;
  int k = m.pin_k[q];
  if (k < 0) return;
  m.pin_cx[q] = m.icx[k] + m.offx[q];
  m.pin_cy[q] = m.icy[k] + m.offy[q];
}

struct IpRec {
  BcState st[2];
  long long hpwl;
};

constexpr int BC_CHUNK = 8;  // grid iterations launched between polls of the done flags

}  // namespace

// One context's InitialPlace run (src/initialPlace.cpp:56-105).
struct RpIpRun {
  rp_ctx* c = nullptr;
  hipStream_t st = nullptr;
  bool host_asm = false;
  int n = 0, np = 0, nn = 0, nparts = 0, eblocks = 0, hblocks = 0;
  int small_r = 0, sell_rows = 0, sell_slices = 0;
  int it = 0;  // IP iterations done
  RpIpAsm A;
  std::vector<int> icx, icy, dlx, dly, hx, hy;
  DevBuf<int> d_dlx, d_dly, d_hx, d_hy, d_mw, d_mh, d_offx, d_offy;
  DevBuf<int4> d_sbox;
  DevBuf<long long> d_hparts;
  DevBuf<float> vec[2][9];  // x (row order), r, r0, v, p, s, t, x and b in sorted order
  DevBuf<double> parts[2];
  DevBuf<IpRec> d_rec;
  IpRec* h_rec = nullptr;
  IpMove mv{};
  BcArgs g{};
  BcState init{};
  SellBufs sb[2];
  SellArgs gs{};
  DevBuf<float> keep[2];
  DevBuf<BcsCarry> carry;
  ~RpIpRun() {
    if (st) (void)hipStreamSynchronize(st);  // no queued work on the buffers released below
    if (h_rec) rp_pinned_free(h_rec);
  }
};

// Tables and pinned staging of a driver solving many runs at once.
struct RpIpBatch {
  DevBuf<SellArgs> d_sell;
  DevBuf<BcArgs> d_bc;
  DevBuf<int2> d_tab, d_htab;
  DevBuf<int> d_done;  // the grid runs' done flags, gathered for one read
  SellArgs* h_sell = nullptr;
  BcArgs* h_bc = nullptr;
  int2* h_tab = nullptr;
  int* h_done = nullptr;
  size_t cap_sell = 0, cap_bc = 0, cap_tab = 0;
  ~RpIpBatch() {
    for (void* p : {(void*)h_sell, (void*)h_bc, (void*)h_tab, (void*)h_done})
      if (p) (void)hipHostFree(p);
  }
};

rp_ctx* rp_ipr_ctx(RpIpRun* r) { return r->c; }

void rp_ipr_free(RpIpRun* r) { delete r; }
RpIpBatch* rp_ipb_create() { return new RpIpBatch(); }
void rp_ipb_free(RpIpBatch* b) { delete b; }

// one workgroup per axis for small systems; RP_IP_SOLVER=blocks keeps the
// grid-wide kernels at every size
bool rp_ip_one_workgroup(int n) {
  const char* solver = getenv("RP_IP_SOLVER");
  return !(solver && strcmp(solver, "blocks") == 0) && bcs_rows_per_thread(n) != 0;
}

int rp_ipr_begin(rp_ctx* c, hipStream_t run_stream, bool host_asm, RpIpRun** out) {
  std::unique_ptr<RpIpRun> R(new RpIpRun());
  R->c = c;
  R->st = run_stream;
  R->host_asm = host_asm;
  hipStream_t st = c->stream;
  const rp_params& p = c->prm;
  const int n = (int)c->place.size(), np = c->n_pin, nn = c->n_net;
  R->n = n;
  R->np = np;
  R->nn = nn;
  RpIpAsm& A = R->A;
  if (int r = rp_ip_asm_init(c, A)) return r;
  // movable state and static per-pin data
  R->icx.resize(n);
  R->icy.resize(n);
  R->dlx.resize(n);
  R->dly.resize(n);
  R->hx.resize(n);
  R->hy.resize(n);
  std::vector<int> mw(n), mh(n);
  for (int j = 0; j < n; j++) {  // row j = movable instance order[j]
    const int k = j, i = c->place[rp_movable_order(c)[j]];
    R->icx[k] = (c->ilx[i] + c->iux[i]) / 2;
    R->icy[k] = (c->ily[i] + c->iuy[i]) / 2;
    R->hx[k] = (c->iux[i] - c->ilx[i]) / 2;
    R->hy[k] = (c->iuy[i] - c->ily[i]) / 2;
    R->dlx[k] = c->dblx[i];
    R->dly[k] = c->dbly[i];
    mw[k] = c->mw[i];
    mh[k] = c->mh[i];
  }
  std::vector<int4> sbox(np);
  for (int q = 0; q < np; q++) {
    int i = c->pin_inst[q];
    if (i >= 0) {
      sbox[q] = make_int4(c->dblx[i], c->dbly[i], c->dblx[i] + c->mw[i], c->dbly[i] + c->mh[i]);
    } else if (c->pin_box[4 * (size_t)q] != INT_MAX) {
      const int* b = &c->pin_box[4 * (size_t)q];
      sbox[q] = make_int4(b[0], b[1], b[2], b[3]);
    } else {
      sbox[q] = make_int4(INT_MAX, INT_MAX, INT_MIN, INT_MIN);
    }
  }
  HIPCHK(c, dupload(R->d_dlx, R->dlx, st));
  HIPCHK(c, dupload(R->d_dly, R->dly, st));
  HIPCHK(c, dupload(R->d_hx, R->hx, st));
  HIPCHK(c, dupload(R->d_hy, R->hy, st));
  HIPCHK(c, dupload(R->d_mw, mw, st));
  HIPCHK(c, dupload(R->d_mh, mh, st));
  HIPCHK(c, dupload(R->d_offx, c->pin_offx, st));
  HIPCHK(c, dupload(R->d_offy, c->pin_offy, st));
  HIPCHK(c, dupload(R->d_sbox, sbox, st));
  R->hblocks = std::max(1, (nn + RP_BLOCK - 1) / RP_BLOCK);
  HIPCHK(c, dalloc(R->d_hparts, R->hblocks));
  HIPCHK(c, hipMemcpyAsync(A.icx.p, R->icx.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(A.icy.p, R->icy.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(A.pin_cx.p, c->pin_cx.data(), sizeof(int) * np, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(A.pin_cy.p, c->pin_cy.data(), sizeof(int) * np, hipMemcpyHostToDevice, st));
  // solver buffers (both axes)
  R->nparts = std::max(1, std::min(BC_MAX_PARTS, (n + GS_WIN - 1) / GS_WIN));  // one block per row window
  R->eblocks = R->nparts;
  for (int ax = 0; ax < 2; ax++) {
    for (auto& v : R->vec[ax]) HIPCHK(c, dalloc(v, n));
    HIPCHK(c, dalloc(R->parts[ax], (size_t)BC_SLOTS * R->nparts));
  }
  HIPCHK(c, dalloc(R->d_rec, 1));
  HIPCHK(c, rp_pinned_alloc((void**)&R->h_rec, sizeof(IpRec)));
  R->mv = IpMove{n, nn, np, A.icx.p, A.icy.p, R->d_dlx.p, R->d_dly.p, R->d_hx.p, R->d_hy.p, R->d_mw.p, R->d_mh.p,
                 A.net_start.p, A.pin_k.p, R->d_offx.p, R->d_offy.p, A.pin_cx.p, A.pin_cy.p, R->d_sbox.p,
                 R->d_hparts.p, R->vec[0][0].p, R->vec[1][0].p};
  BcArgs& g = R->g;
  g.n = n;
  g.nparts = R->nparts;
  for (int ax = 0; ax < 2; ax++) {
    BcAxis& a = g.ax[ax];
    a.xo = R->vec[ax][0].p;
    a.x = R->vec[ax][7].p;
    a.r = R->vec[ax][1].p;
    a.r0 = R->vec[ax][2].p;
    a.v = R->vec[ax][3].p;
    a.p = R->vec[ax][4].p;
    a.s = R->vec[ax][5].p;
    a.t = R->vec[ax][6].p;
    a.parts = R->parts[ax].p;
    a.st = &R->d_rec.p->st[ax];
  }
  const float tol = 1.1920928955078125e-07f;  // NumTraits<float>::epsilon()
  R->init.tol2 = (double)tol * (double)tol;
  R->init.eps2 = 1.1920928955078125e-07 * 1.1920928955078125e-07;
  R->init.max_iter = p.initial_place_max_solver_iter;
  R->small_r = rp_ip_one_workgroup(n) ? bcs_rows_per_thread(n) : 0;
  R->sell_rows = BCS_THREADS * R->small_r;
  R->sell_slices = R->sell_rows / SELL_C;
  if (R->small_r) {
    HIPCHK(c, dalloc(R->carry, 2));
    for (int ax = 0; ax < 2; ax++) {
      SellBufs& B = R->sb[ax];
      HIPCHK(c, dalloc(R->keep[ax], 5 * (size_t)n));
      HIPCHK(c, dalloc(B.key, n));
      HIPCHK(c, dalloc(B.key2, n));
      HIPCHK(c, dalloc(B.id, n));
      HIPCHK(c, dalloc(B.perm, n));
      HIPCHK(c, dalloc(B.iperm, n));
      HIPCHK(c, dalloc(B.slen, n));
      HIPCHK(c, dalloc(B.wslots, R->sell_slices));
      HIPCHK(c, dalloc(B.soff, R->sell_slices));
      size_t t1 = 0, t2 = 0;
      HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, t1, B.key.p, B.key2.p, B.id.p, B.perm.p, n, 0, 16,
                                                   st));
      HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, t2, B.wslots.p, B.soff.p, R->sell_slices, st));
      HIPCHK(c, dalloc(B.temp, std::max(t1, t2)));
    }
  }
  if (!R->small_r) {
    const int nwin = (n + GS_WIN - 1) / GS_WIN;
    for (int ax = 0; ax < 2; ax++) {
      SellBufs& B = R->sb[ax];
      HIPCHK(c, dalloc(B.perm, n));
      HIPCHK(c, dalloc(B.iperm, n));
      HIPCHK(c, dalloc(B.slen, n));
      HIPCHK(c, dalloc(B.qslice, n));
      HIPCHK(c, dalloc(B.sinfo, (size_t)nwin * GS_SMAX));
      HIPCHK(c, dalloc(B.nsl, nwin));
      HIPCHK(c, dalloc(B.wtot, nwin));
    }
  }
  HIPCHK(c, hipStreamSynchronize(st));
  *out = R.release();
  return 0;
}

int rp_ipr_assemble_a(RpIpRun* R) {
  if (R->host_asm) return 0;
  return rp_ip_assemble_a(R->c, R->A, R->st);
}

int rp_ipr_assemble_b(RpIpRun* R) {
  rp_ctx* c = R->c;
  hipStream_t st = R->st;
  RpIpAsm& A = R->A;
  const int n = R->n;
  const auto ta = std::chrono::steady_clock::now();
  if (R->host_asm) {
    if (int r = rp_ip_host_assemble(c, A)) return r;
  } else {
    if (int r = rp_ip_assemble_b(c, A, st)) return r;
  }
  const double t_asm = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
  if (t_asm > 20.0 && getenv("RP_IP_TIMING"))
    fprintf(stderr, "[rp] slow assembly enqueue: %.1f ms (%d rows)\n", t_asm, n);
  BcArgs& g = R->g;
  for (int ax = 0; ax < 2; ax++) {
    BcAxis& a = g.ax[ax];
    a.rowp = A.rowp[ax].p;
    a.col = A.col[ax].p;
    a.val = A.val[ax].p;
    a.bo = A.b[ax].p;
    a.b = R->vec[ax][8].p;
  }
  IpRec r0{};
  r0.st[0] = r0.st[1] = R->init;
  *R->h_rec = r0;
  HIPCHK(c, hipMemcpyAsync(R->d_rec.p, R->h_rec, sizeof(IpRec), hipMemcpyHostToDevice, st));
  const dim3 gn((n + RP_BLOCK - 1) / RP_BLOCK), blk(RP_BLOCK);
  hipLaunchKernelGGL(k_ip_warm, gn, blk, 0, st, R->mv);
  if (!R->small_r) {  // sliced-ELL copy of each axis' matrix for the grid-wide solve
    const int nwin = (n + GS_WIN - 1) / GS_WIN;
    for (int ax = 0; ax < 2; ax++) {
      SellBufs& B = R->sb[ax];
      // slots: within a window, the one-lane slices hold at most the
      // nonzeros, 64 x 19 for the first one and 3 rounding slots a row; a full
      // slice of g >= 2 lanes a row holds at most twice its nonzeros, and a
      // window has at most six partial ones of 1024 slots
      const size_t cap = 2 * A.col[ax].n + 3 * (size_t)n + 8192 * (size_t)nwin;
      if (cap >= (size_t)INT_MAX) return fail(c, RP_ERR_ARG, "initial place: matrix too large for 32-bit offsets");
      if (B.sval.n < cap) {
        HIPCHK(c, dalloc(B.sval, cap));
        HIPCHK(c, dalloc(B.scol32, cap));
      }
      BcAxis& a = g.ax[ax];
      hipLaunchKernelGGL(k_gsell_order, dim3(nwin), dim3(GS_WIN), 0, st, (const int*)A.rowp[ax].p, n, B.perm.p,
                         B.iperm.p, B.slen.p, B.qslice.p, B.sinfo.p, B.nsl.p, B.wtot.p);
      hipLaunchKernelGGL(k_gsell_fill, dim3(nwin), dim3(GS_WIN), 0, st, (const int*)A.rowp[ax].p,
                         (const int*)A.col[ax].p, (const float*)A.val[ax].p, (const float*)A.b[ax].p,
                         (const float*)a.xo, (const int*)B.perm.p, (const int*)B.iperm.p, (const int*)B.slen.p,
                         (const int*)B.qslice.p, (const int*)B.nsl.p, (const int*)B.wtot.p, n, B.sinfo.p,
                         B.sval.p, B.scol32.p, R->vec[ax][8].p, a.x);
      HIPCHK(c, hipGetLastError());
      a.perm = B.perm.p;
      a.slen = B.slen.p;
      a.sinfo = B.sinfo.p;
      a.nsl = B.nsl.p;
      a.sval = B.sval.p;
      a.scol = B.scol32.p;
    }
    return 0;
  }
  // sorted SELL copy of each axis' matrix
  for (int ax = 0; ax < 2; ax++) {
    SellBufs& B = R->sb[ax];
    // nnz + one slice of the longest row + the rounding of every slice
    const size_t cap = A.col[ax].n + (size_t)SELL_C * n + (size_t)SELL_C * (SELL_PAD - 1) * R->sell_slices;
    if (B.sval.n < cap) {
      HIPCHK(c, dalloc(B.sval, cap));
      HIPCHK(c, dalloc(B.scol, cap));
    }
    hipLaunchKernelGGL(k_sell_keys, gn, blk, 0, st, (const int*)A.rowp[ax].p, n, B.key.p, B.id.p);
    size_t tb = B.temp.n;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(B.temp.p, tb, B.key.p, B.key2.p, B.id.p, B.perm.p, n, 0, 16, st));
    hipLaunchKernelGGL(k_sell_meta, grid_of(std::max(n, R->sell_slices)), blk, 0, st, (const int*)A.rowp[ax].p,
                       (const int*)B.perm.p, n, R->sell_slices, B.iperm.p, B.slen.p, B.wslots.p);
    tb = B.temp.n;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(B.temp.p, tb, B.wslots.p, B.soff.p, R->sell_slices, st));
    hipLaunchKernelGGL(k_sell_fill, grid_of(R->sell_rows), blk, 0, st, (const int*)A.rowp[ax].p,
                       (const int*)A.col[ax].p, (const float*)A.val[ax].p, (const int*)B.perm.p,
                       (const int*)B.iperm.p, (const int*)B.soff.p, n, R->sell_rows, B.sval.p, B.scol.p);
    HIPCHK(c, hipGetLastError());
    SellAxis& a = R->gs.ax[ax];
    a.rowp = g.ax[ax].rowp;
    a.col = g.ax[ax].col;
    a.val = g.ax[ax].val;
    a.b = g.ax[ax].bo;
    a.x = g.ax[ax].xo;
    a.perm = B.perm.p;
    a.slen = B.slen.p;
    a.width = B.wslots.p;
    a.soff = B.soff.p;
    a.sval = B.sval.p;
    a.scol = B.scol.p;
    a.st = g.ax[ax].st;
    a.keep = R->keep[ax].p;
    a.carry = R->carry.p + ax;
  }
  R->gs.n = n;
  return 0;
}

namespace {

template <typename T>
int grow_pinned(rp_ctx* c, T*& h, DevBuf<T>& d, size_t& cap, size_t need) {
  if (cap >= need && h) return 0;
  size_t n = std::max<size_t>(need, 2 * cap);
  if (h) (void)hipHostFree(h);
  h = nullptr;
  HIPCHK(c, hipHostMalloc((void**)&h, n * sizeof(T), hipHostMallocDefault));
  HIPCHK(c, dalloc(d, n));
  cap = n;
  return 0;
}

int solve_guard(const RpIpRun* R) { return 2 * std::max(0, R->init.max_iter) + 2; }

void launch_small(int r, dim3 grid, size_t lds, hipStream_t st, const SellArgs* gs, int first) {
  switch (r) {
    case 1: hipLaunchKernelGGL(kb_bc_small<1>, grid, dim3(BCS_THREADS), lds, st, gs, first, BCS_PASSES); break;
    case 2: hipLaunchKernelGGL(kb_bc_small<2>, grid, dim3(BCS_THREADS), lds, st, gs, first, BCS_PASSES); break;
    case 4: hipLaunchKernelGGL(kb_bc_small<4>, grid, dim3(BCS_THREADS), lds, st, gs, first, BCS_PASSES); break;
    default: hipLaunchKernelGGL(kb_bc_small<8>, grid, dim3(BCS_THREADS), lds, st, gs, first, BCS_PASSES); break;
  }
}

void launch_small1(int r, size_t lds, hipStream_t st, const SellArgs& g, int first) {
  switch (r) {
    case 1: hipLaunchKernelGGL(k_bc_small<1>, dim3(2), dim3(BCS_THREADS), lds, st, g, first, BCS_PASSES); break;
    case 2: hipLaunchKernelGGL(k_bc_small<2>, dim3(2), dim3(BCS_THREADS), lds, st, g, first, BCS_PASSES); break;
    case 4: hipLaunchKernelGGL(k_bc_small<4>, dim3(2), dim3(BCS_THREADS), lds, st, g, first, BCS_PASSES); break;
    default: hipLaunchKernelGGL(k_bc_small<8>, dim3(2), dim3(BCS_THREADS), lds, st, g, first, BCS_PASSES); break;
  }
}

}  // namespace

// The solves of k runs.  One-workgroup runs of one row class share a launch
// (2 workgroups per run); grid runs share each of the seven launches through
// a (run, block) table, each run keeping its own block count, so every run
// computes exactly what it computes alone.
int rp_ipr_solve(RpIpRun* const* runs, int k, hipStream_t st, RpIpBatch* bt) {
  if (k <= 0) return 0;
  rp_ctx* c0 = runs[0]->c;
  const dim3 blk(RP_BLOCK), gblk(GS_WIN);
  if (k == 1 || !bt) {  // one run: launches of its own
    for (int j = 0; j < k; j++) {
      RpIpRun* R = runs[j];
      rp_ctx* c = R->c;
      if (R->small_r) {
        const size_t lds = 4 * sizeof(float) * (size_t)R->n;
        for (int first = 0; first < solve_guard(R); first += BCS_PASSES) launch_small1(R->small_r, lds, st, R->gs, first);
        HIPCHK(c, hipGetLastError());
        continue;
      }
      const BcArgs& g = R->g;
      const dim3 gb(R->nparts, 2), ge(R->eblocks, 2), g1(1, 2);
      const int cap = 2 * std::max(0, R->init.max_iter) + 1;  // i restarts from 0 at most once
      hipLaunchKernelGGL(k_bc_init, gb, gblk, 0, st, g);
      int done_iters = 0;
      int* h_done = &R->h_rec->st[0].done;  // scratch: the record is rewritten by finish_a
      for (;;) {
        for (int i = 0; i < BC_CHUNK; i++) {
          hipLaunchKernelGGL(k_bc_head, g1, dim3(1024), 0, st, g);
          hipLaunchKernelGGL(k_bc_restart, gb, gblk, 0, st, g);
          hipLaunchKernelGGL(k_bc_p, ge, blk, 0, st, g);
          hipLaunchKernelGGL(k_bc_v, gb, gblk, 0, st, g);
          hipLaunchKernelGGL(k_bc_s, ge, blk, 0, st, g);
          hipLaunchKernelGGL(k_bc_t, gb, gblk, 0, st, g);
          hipLaunchKernelGGL(k_bc_xr, ge, blk, 0, st, g);
        }
        done_iters += BC_CHUNK;
        // no solve ends before max_iter iterations unless it converges to
        // float epsilon, so the first poll waits for them (the kernels of a
        // finished solve return at once; polling changes no result)
        if (done_iters < std::min(cap, R->init.max_iter)) continue;
        hipLaunchKernelGGL(k_bc_head, g1, dim3(1024), 0, st, g);  // closes the chunk's last iteration
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(&h_done[0], &R->d_rec.p->st[0].done, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipMemcpyAsync(&R->h_rec->st[1].done, &R->d_rec.p->st[1].done, sizeof(int),
                                 hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if ((R->h_rec->st[0].done && R->h_rec->st[1].done) || done_iters >= cap) break;
      }
      hipLaunchKernelGGL(k_bc_unperm, ge, blk, 0, st, g);
      HIPCHK(c, hipGetLastError());
    }
    return 0;
  }
  // one-workgroup runs, one launch set per row class
  std::vector<RpIpRun*> small[BCS_MAX_R + 1], grid;
  for (int j = 0; j < k; j++) {
    if (runs[j]->small_r)
      small[runs[j]->small_r].push_back(runs[j]);
    else
      grid.push_back(runs[j]);
  }
  size_t nsmall = 0;
  for (auto& v : small) nsmall += v.size();
  if (nsmall) {
    if (int r = grow_pinned(c0, bt->h_sell, bt->d_sell, bt->cap_sell, nsmall)) return r;
    size_t at = 0;
    for (int rc = 1; rc <= BCS_MAX_R; rc++) {
      if (small[rc].empty()) continue;
      int nmax = 0, guard = 0;
      for (size_t j = 0; j < small[rc].size(); j++) {
        bt->h_sell[at + j] = small[rc][j]->gs;
        nmax = std::max(nmax, small[rc][j]->n);
        guard = std::max(guard, solve_guard(small[rc][j]));
      }
      HIPCHK(c0, hipMemcpyAsync(bt->d_sell.p + at, bt->h_sell + at, sizeof(SellArgs) * small[rc].size(),
                                hipMemcpyHostToDevice, st));
      const size_t lds = 4 * sizeof(float) * (size_t)nmax;
      for (int first = 0; first < guard; first += BCS_PASSES)
        launch_small(rc, dim3(2 * (unsigned)small[rc].size()), lds, st, bt->d_sell.p + at, first);
      HIPCHK(c0, hipGetLastError());
      at += small[rc].size();
    }
  }
  if (grid.empty()) return 0;
  // grid runs: (run, block) tables for the group-per-row kernels (nparts
  // blocks a run) and the thread-per-row kernels (eblocks), (run, 0) for the head
  const int ng = (int)grid.size();
  size_t ntab = 0, netab = 0;
  int cap = 0, first_poll = INT_MAX;
  for (RpIpRun* R : grid) {
    ntab += R->nparts;
    netab += R->eblocks;
    cap = std::max(cap, 2 * std::max(0, R->init.max_iter) + 1);
    first_poll = std::min(first_poll, R->init.max_iter);
  }
  if (int r = grow_pinned(c0, bt->h_bc, bt->d_bc, bt->cap_bc, ng)) return r;
  if (int r = grow_pinned(c0, bt->h_tab, bt->d_tab, bt->cap_tab, ntab + ng + netab)) return r;
  if (!bt->h_done) HIPCHK(c0, hipHostMalloc((void**)&bt->h_done, sizeof(int) * 2 * 4096, hipHostMallocDefault));
  if (!bt->d_done.p) HIPCHK(c0, dalloc(bt->d_done, 2 * 4096));
  if (ng > 4096) return fail(c0, RP_ERR_ARG, "initial place: more than 4096 grid solves in one batch");
  size_t t = 0;
  for (int j = 0; j < ng; j++) {
    bt->h_bc[j] = grid[j]->g;
    for (int b = 0; b < grid[j]->nparts; b++) bt->h_tab[t++] = make_int2(j, b);
  }
  for (int j = 0; j < ng; j++) bt->h_tab[t++] = make_int2(j, 0);
  for (int j = 0; j < ng; j++)
    for (int b = 0; b < grid[j]->eblocks; b++) bt->h_tab[t++] = make_int2(j, b);
  HIPCHK(c0, hipMemcpyAsync(bt->d_bc.p, bt->h_bc, sizeof(BcArgs) * ng, hipMemcpyHostToDevice, st));
  HIPCHK(c0, hipMemcpyAsync(bt->d_tab.p, bt->h_tab, sizeof(int2) * t, hipMemcpyHostToDevice, st));
  const BcArgs* gsd = bt->d_bc.p;
  const int2* tab = bt->d_tab.p;
  const int2* htab = bt->d_tab.p + ntab;
  const int2* etab = bt->d_tab.p + ntab + ng;
  const dim3 gb((unsigned)ntab, 2), ge((unsigned)netab, 2), g1((unsigned)ng, 2);
  // RP_IP_TIMING: event times of the launches of one iteration (the tenth)
  // and of the whole solve
  const bool timing = getenv("RP_IP_TIMING") != nullptr;
  hipEvent_t ev[10] = {};
  if (timing) {
    for (auto& e : ev) (void)hipEventCreate(&e);
    (void)hipEventRecord(ev[8], st);
  }
  hipLaunchKernelGGL(kb_bc_init, gb, gblk, 0, st, gsd, tab);
  int done_iters = 0, polls = 0;
  const auto th0 = std::chrono::steady_clock::now();
  double t_enq = 0.0;
  for (;;) {
    for (int i = 0; i < BC_CHUNK; i++) {
      const bool tk = timing && done_iters + i == 9;
      if (tk) (void)hipEventRecord(ev[0], st);
      hipLaunchKernelGGL(kb_bc_head, g1, dim3(1024), 0, st, gsd, htab);
      if (tk) (void)hipEventRecord(ev[1], st);
      hipLaunchKernelGGL(kb_bc_restart, gb, gblk, 0, st, gsd, tab);
      if (tk) (void)hipEventRecord(ev[2], st);
      hipLaunchKernelGGL(kb_bc_p, ge, blk, 0, st, gsd, etab);
      if (tk) (void)hipEventRecord(ev[3], st);
      hipLaunchKernelGGL(kb_bc_v, gb, gblk, 0, st, gsd, tab);
      if (tk) (void)hipEventRecord(ev[4], st);
      hipLaunchKernelGGL(kb_bc_s, ge, blk, 0, st, gsd, etab);
      if (tk) (void)hipEventRecord(ev[5], st);
      hipLaunchKernelGGL(kb_bc_t, gb, gblk, 0, st, gsd, tab);
      if (tk) (void)hipEventRecord(ev[6], st);
      hipLaunchKernelGGL(kb_bc_xr, ge, blk, 0, st, gsd, etab);
      if (tk) (void)hipEventRecord(ev[7], st);
    }
    done_iters += BC_CHUNK;
    if (done_iters < std::min(cap, first_poll)) continue;  // as for one run
    if (!polls) t_enq = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count();
    hipLaunchKernelGGL(kb_bc_head, g1, dim3(1024), 0, st, gsd, htab);
    HIPCHK(c0, hipGetLastError());
    hipLaunchKernelGGL(k_bc_done_gather, dim3((2 * ng + RP_BLOCK - 1) / RP_BLOCK), blk, 0, st, gsd, ng,
                       bt->d_done.p);
    HIPCHK(c0, hipMemcpyAsync(bt->h_done, bt->d_done.p, sizeof(int) * 2 * ng, hipMemcpyDeviceToHost, st));
    HIPCHK(c0, hipStreamSynchronize(st));
    polls++;
    bool all = true;
    for (int j = 0; j < 2 * ng; j++) all = all && bt->h_done[j];
    if (all || done_iters >= cap) break;
  }
  hipLaunchKernelGGL(kb_bc_unperm, ge, blk, 0, st, gsd, etab);
  if (timing) {
    (void)hipEventRecord(ev[9], st);
    (void)hipEventSynchronize(ev[9]);
    float t[8] = {};
    for (int k = 0; k < 7; k++) (void)hipEventElapsedTime(&t[k], ev[k], ev[k + 1]);
    (void)hipEventElapsedTime(&t[7], ev[8], ev[9]);
    fprintf(stderr, "[rp] grid solves of %d runs (%zu blocks), %d iterations, %d polls, %.2f ms on the stream, "
            "%.2f ms to enqueue the first %d: head %.1f restart %.1f p %.1f v %.1f s %.1f t %.1f xr %.1f us\n", ng,
            ntab, done_iters, polls, t[7], t_enq, std::min(cap, first_poll),
            1e3 * t[0], 1e3 * t[1], 1e3 * t[2], 1e3 * t[3], 1e3 * t[4], 1e3 * t[5], 1e3 * t[6]);
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  HIPCHK(c0, hipGetLastError());
  return 0;
}

int rp_ipr_finish_a(RpIpRun* R) {
  rp_ctx* c = R->c;
  hipStream_t st = R->st;
  const dim3 blk(RP_BLOCK), gn((R->n + RP_BLOCK - 1) / RP_BLOCK), gp((R->np + RP_BLOCK - 1) / RP_BLOCK);
  // log record: HPWL before updateCoordi (src/initialPlace.cpp:88-91)
  hipLaunchKernelGGL(k_ip_hpwl, dim3(R->hblocks), blk, 0, st, R->mv);
  hipLaunchKernelGGL(k_ip_hpwl_final, dim3(1), dim3(1024), 0, st, (const long long*)R->d_hparts.p, R->hblocks,
                     &R->d_rec.p->hpwl);
  hipLaunchKernelGGL(k_ip_update, gn, blk, 0, st, R->mv);
  hipLaunchKernelGGL(k_ip_pins, gp, blk, 0, st, R->mv);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(R->h_rec, R->d_rec.p, sizeof(IpRec), hipMemcpyDeviceToHost, st));
  return 0;
}

int rp_ipr_finish_b(RpIpRun* R, bool* stop) {
  rp_ctx* c = R->c;
  const IpRec& rec = *R->h_rec;
  R->it++;
  if (!rec.st[0].done || !rec.st[1].done) return fail(c, RP_ERR_STATE, "initial place: BiCGSTAB did not terminate");
  const float err = std::max(rec.st[0].err, rec.st[1].err);
  if (R->small_r && getenv("RP_IP_TIMING"))
    fprintf(stderr, "[rp] initial place %d: one-workgroup solves of %d rows, %d / %d iterations\n", R->it, R->n,
            rec.st[0].pending, rec.st[1].pending);
  c->ip_err.push_back(err);
  c->ip_hpwl.push_back(rec.hpwl);
  *stop = (err <= 1e-5 && R->it >= 5) || R->it >= c->prm.initial_place_max_iter;
  return 0;
}

int rp_ipr_end(RpIpRun* R) {
  rp_ctx* c = R->c;
  hipStream_t st = R->st;
  const int n = R->n, np = R->np;
  // final positions back to the host model
  HIPCHK(c, hipMemcpyAsync(R->icx.data(), R->A.icx.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(R->icy.data(), R->A.icy.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(R->dlx.data(), R->d_dlx.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(R->dly.data(), R->d_dly.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  const std::vector<int>& order = rp_movable_order(c);
  for (int k = 0; k < n; k++) {
    int i = c->place[order[k]];
    int h0 = R->hx[k], h1 = R->hy[k];
    c->ilx[i] = R->icx[k] - h0;
    c->iux[i] = R->icx[k] + h0;
    c->ily[i] = R->icy[k] - h1;
    c->iuy[i] = R->icy[k] + h1;
    c->dblx[i] = R->dlx[k];
    c->dbly[i] = R->dly[k];
  }
  for (int q = 0; q < np; q++) {
    int i = c->pin_inst[q];
    if (i < 0 || c->inst2place[i] < 0) continue;
    c->pin_cx[q] = (c->ilx[i] + c->iux[i]) / 2 + c->pin_offx[q];
    c->pin_cy[q] = (c->ily[i] + c->iuy[i]) / 2 + c->pin_offy[q];
  }
  return 0;
}

// The InitialPlace loop (src/initialPlace.cpp:56-105) of one context with
// every iteration on the device; the host reads one record (errors, HPWL)
// per iteration.
int rp_ip_device_loop(rp_ctx* c, bool host_asm, double* t_asm, double* t_solve) {
  using clk = std::chrono::steady_clock;
  auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
  hipStream_t st = c->stream;
  RpIpRun* raw = nullptr;
  if (int r = rp_ipr_begin(c, st, host_asm, &raw)) return r;
  std::unique_ptr<RpIpRun> R(raw);
  for (;;) {
    auto ta = clk::now();
    if (int r = rp_ipr_assemble_a(R.get())) return r;
    if (!host_asm) HIPCHK(c, hipStreamSynchronize(st));
    if (int r = rp_ipr_assemble_b(R.get())) return r;
    *t_asm += since(ta);
    auto ts = clk::now();
    RpIpRun* one = R.get();
    if (int r = rp_ipr_solve(&one, 1, st, nullptr)) return r;
    if (int r = rp_ipr_finish_a(R.get())) return r;
    HIPCHK(c, hipStreamSynchronize(st));
    bool stop = false;
    if (int r = rp_ipr_finish_b(R.get(), &stop)) return r;
    *t_solve += since(ts);
    if (stop) break;
  }
  return rp_ipr_end(R.get());
}

// Loads this file's code object (its kernels and the hipCUB ones it
// instantiates) on the current device.  The placement service calls it once
// before it starts its set-up threads, so the first launches of many threads
// at once do not each trigger the load.
hipError_t rp_ip_load_code() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_ip_segsum));
}
