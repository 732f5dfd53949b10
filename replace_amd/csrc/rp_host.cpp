// rp_host.cpp -- host side of the C ABI (include/replace/rp_capi.h).
//
// One-time work runs on the host (it is O(N) and cheap next to the loop):
//   PlacerBase::init derivations         (reference src/placerBase.cpp:605-846)
//   B2B matrix assembly of InitialPlace  (src/initialPlace.cpp:128-378)
//   NesterovBase::init: fillers, bins,
//     density sizes                      (src/nesterovBase.cpp:631-741, 916-1155)
// Everything per iteration runs on the device: BiCGSTAB (SpMV + FP64 dots),
// and the whole Nesterov loop (gradient evaluations, back-tracking, schedules
// and the stop test) driven by the controller kernel in rp_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/replace/rp_capi.h"
#include "rp_device.h"

#include "rp_dev_struct.h"


#include "rp_ctx.h"

namespace {
constexpr long double kSqrt2 = 1.414213562373095048801L;  // src/nesterovBase.cpp:13
constexpr long double kPi = 3.141592653589793238462L;     // src/fft.cpp:9
}  // namespace

static inline int idiv_upper(int r, int s) {
  // ceil-like bin upper index of getMinMaxIdx / getDensityMinMaxIdx
  int fm = r >= s ? r % s : r;
  return (fm == 0) ? r / s : r / s + 1;
}

// ===========================================================================
// PlacerBase
// ===========================================================================
namespace {
std::mutex g_pool_mu;
bool g_pool_ready[64];
}  // namespace

hipError_t rp_dev_malloc(void** p, size_t bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev >= 0 && dev < 64) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (!g_pool_ready[dev]) {
      // keep freed pool memory for later allocations instead of returning it
      hipMemPool_t pool;
      if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      }
      g_pool_ready[dev] = true;
    }
  }
  return hipMallocAsync(p, bytes, nullptr);
}

void rp_dev_free(void* p) { (void)hipFreeAsync(p, nullptr); }

// Streams of placement contexts are kept for reuse once a context is
// released: creating and destroying a stream per job costs milliseconds and,
// in a service, stalls the other streams' work while it runs.
namespace {
std::mutex g_stream_mu;
std::vector<hipStream_t> g_free_streams[64];
}  // namespace

hipError_t rp_stream_acquire(int dev, hipStream_t* st) {
  if (dev >= 0 && dev < 64) {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if (!g_free_streams[dev].empty()) {
      *st = g_free_streams[dev].back();
      g_free_streams[dev].pop_back();
      return hipSuccess;
    }
  }
  return hipStreamCreateWithFlags(st, hipStreamNonBlocking);
}

// the stream must be idle (its work synchronised) when it is handed back
void rp_stream_release(int dev, hipStream_t st) {
  if (!st) return;
  if (dev >= 0 && dev < 64) {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    g_free_streams[dev].push_back(st);
    return;
  }
  (void)hipStreamDestroy(st);
}

namespace {
std::mutex g_pin_mu;
std::vector<void*> g_pin_free[48];  // cached blocks by size class (2^k bytes)
std::unordered_map<void*, int> g_pin_class;
int pin_class(size_t bytes) {
  int k = 8;  // 256 bytes at least
  while (((size_t)1 << k) < bytes) k++;
  return k;
}
}  // namespace

hipError_t rp_pinned_alloc(void** p, size_t bytes) {
  const int k = pin_class(bytes);
  if (k >= 48) return hipErrorInvalidValue;
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    if (!g_pin_free[k].empty()) {
      *p = g_pin_free[k].back();
      g_pin_free[k].pop_back();
      return hipSuccess;
    }
  }
  hipError_t e = hipHostMalloc(p, (size_t)1 << k, hipHostMallocDefault);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_pin_mu);
  g_pin_class[*p] = k;
  return hipSuccess;
}

void rp_pinned_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pin_mu);
  auto it = g_pin_class.find(p);
  if (it != g_pin_class.end()) g_pin_free[it->second].push_back(p);
}

extern "C" void rp_params_default(rp_params* p) {
  // Replace::Replace, src/replace.cpp:13-33 and NesterovPlaceVars / NesterovBaseVars constants
  p->initial_place_max_iter = 20;
  p->initial_place_min_diff_length = 1500;
  p->initial_place_max_solver_iter = 100;
  p->initial_place_max_fanout = 200;
  p->initial_place_net_weight_scale = 800.0f;
  p->nesterov_place_max_iter = 2000;
  p->bin_grid_cnt_x = 0;
  p->bin_grid_cnt_y = 0;
  p->target_overflow = 0.1f;
  p->target_density = 1.0f;
  p->init_density_penalty = 0.00008f;
  p->init_wire_length_coef = 0.25f;
  p->min_phi_coef = 0.95f;
  p->max_phi_coef = 1.05f;
  p->reference_hpwl = 446000000.0f;
  p->incremental_place_mode = 0;
  p->verbose = 0;
  p->max_back_track = 10;
  p->min_preconditioner = 1.0f;
  p->initial_prev_coordi_update_coef = 100.0f;
  p->min_wire_length_force_bar = -300.0f;
  p->poisson_transform = 0;
}

static int inst_cx(const rp_ctx* c, int i) { return (c->ilx[i] + c->iux[i]) / 2; }
static int inst_cy(const rp_ctx* c, int i) { return (c->ily[i] + c->iuy[i]) / 2; }

static void inst_update_pins(rp_ctx* c, int i) {
  // Pin::updateLocation, src/placerBase.cpp:408-412
  for (int k = c->ipin_start[i]; k < c->ipin_start[i + 1]; k++) {
    int q = c->ipins[k];
    c->pin_cx[q] = inst_cx(c, i) + c->pin_offx[q];
    c->pin_cy[q] = inst_cy(c, i) + c->pin_offy[q];
  }
}

static void inst_set_center(rp_ctx* c, int i, int x, int y) {
  // Instance::setCenterLocation, src/placerBase.cpp:127-140 (odd sizes shrink by one)
  int hx = (c->iux[i] - c->ilx[i]) / 2, hy = (c->iuy[i] - c->ily[i]) / 2;
  c->ilx[i] = x - hx;
  c->ily[i] = y - hy;
  c->iux[i] = x + hx;
  c->iuy[i] = y + hy;
  inst_update_pins(c, i);
}

static std::pair<int, int> site_range(int ll, int uu, int core, int size, int lo, int hi) {
  // getMinMaxIdx, src/placerBase.cpp:984-993
  int a = (ll - core) / size;
  int b = idiv_upper(uu - core, size);
  return {std::max(lo, a), std::min(hi, b)};
}

extern "C" int rp_create(const rp_netlist* nl, const rp_params* prm, rp_ctx** out) {
  if (!out) return RP_ERR_ARG;
  *out = nullptr;
  if (!nl || !prm || nl->num_rows <= 0 || nl->site_w <= 0 || nl->site_h <= 0) return RP_ERR_ARG;
  std::unique_ptr<rp_ctx> c(new rp_ctx());
  c->prm = *prm;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    *out = c.release();
    return fail(*out, RP_ERR_NO_DEVICE, "no HIP device visible");
  }
  (void)hipGetDevice(&c->device);
  if (rp_stream_acquire(c->device, &c->stream) != hipSuccess) {
    *out = c.release();
    return fail(*out, RP_ERR_HIP, "hipStreamCreate failed");
  }
  if (rp_pinned_alloc((void**)&c->h_state, sizeof(RpState)) != hipSuccess) {
    *out = c.release();
    return fail(*out, RP_ERR_HIP, "hipHostMalloc failed");
  }
  // core = bbox of rows (src/placerBase.cpp:961-976)
  int minX = INT_MAX, minY = INT_MAX, maxX = INT_MIN, maxY = INT_MIN;
  for (int r = 0; r < nl->num_rows; r++) {
    minX = std::min(nl->row_box[4 * r + 0], minX);
    minY = std::min(nl->row_box[4 * r + 1], minY);
    maxX = std::max(nl->row_box[4 * r + 2], maxX);
    maxY = std::max(nl->row_box[4 * r + 3], maxY);
  }
  c->core_lx = minX;
  c->core_ly = minY;
  c->core_ux = maxX;
  c->core_uy = maxY;
  c->site_w = nl->site_w;
  c->site_h = nl->site_h;
  c->die_lx = nl->die_lx;
  c->die_ly = nl->die_ly;
  c->die_ux = nl->die_ux;
  c->die_uy = nl->die_uy;
  c->dbu = nl->dbu_per_micron;
  const int nr = nl->num_insts;
  // fragmented-row dummy instances (src/placerBase.cpp:760-846)
  int scx = (c->core_ux - c->core_lx) / c->site_w;
  int scy = (c->core_uy - c->core_ly) / c->site_h;
  std::vector<uint8_t> grid((size_t)std::max(scx, 0) * std::max(scy, 0), 0);
  for (int r = 0; r < nl->num_rows; r++) {
    auto px = site_range(nl->row_box[4 * r], nl->row_box[4 * r + 2], c->core_lx, c->site_w, 0, scx);
    auto py = site_range(nl->row_box[4 * r + 1], nl->row_box[4 * r + 3], c->core_ly, c->site_h, 0, scy);
    for (int i = px.first; i < px.second; i++)
      for (int j = py.first; j < py.second; j++) grid[(size_t)j * scx + i] = 1;
  }
  for (int i = 0; i < nr; i++) {
    if (!nl->inst_fixed[i]) continue;
    auto px = site_range(nl->inst_box[4 * i], nl->inst_box[4 * i + 2], c->core_lx, c->site_w, 0, scx);
    auto py = site_range(nl->inst_box[4 * i + 1], nl->inst_box[4 * i + 3], c->core_ly, c->site_h, 0, scy);
    for (int a = px.first; a < px.second; a++)
      for (int b = py.first; b < py.second; b++) grid[(size_t)b * scx + a] = 2;
  }
  std::vector<std::array<int, 4>> dummies;
  for (int j = 0; j < scy; j++) {
    for (int i = 0; i < scx; i++) {
      if (grid[(size_t)j * scx + i] != 0) continue;
      int sx = i;
      while (i < scx && grid[(size_t)j * scx + i] == 0) i++;
      dummies.push_back({c->core_lx + c->site_w * sx, c->core_ly + c->site_h * j, c->core_lx + c->site_w * i,
                         c->core_ly + c->site_h * (j + 1)});
    }
  }
  grid.clear();
  grid.shrink_to_fit();
  c->n_real = nr;
  int ni = nr + (int)dummies.size();
  c->ilx.resize(ni);
  c->ily.resize(ni);
  c->iux.resize(ni);
  c->iuy.resize(ni);
  c->mw.assign(ni, 0);
  c->mh.assign(ni, 0);
  c->dblx.assign(ni, 0);
  c->dbly.assign(ni, 0);
  c->fixed.assign(ni, 1);
  c->inst2place.assign(ni, -1);
  for (int i = 0; i < nr; i++) {
    c->ilx[i] = nl->inst_box[4 * i];
    c->ily[i] = nl->inst_box[4 * i + 1];
    c->iux[i] = nl->inst_box[4 * i + 2];
    c->iuy[i] = nl->inst_box[4 * i + 3];
    c->mw[i] = c->iux[i] - c->ilx[i];
    c->mh[i] = c->iuy[i] - c->ily[i];
    c->dblx[i] = c->ilx[i];
    c->dbly[i] = c->ily[i];
    c->fixed[i] = nl->inst_fixed[i] ? 1 : 0;
  }
  for (size_t k = 0; k < dummies.size(); k++) {
    int i = nr + (int)k;
    c->ilx[i] = dummies[k][0];
    c->ily[i] = dummies[k][1];
    c->iux[i] = dummies[k][2];
    c->iuy[i] = dummies[k][3];
  }
  // classification and areas (src/placerBase.cpp:639-680)
  for (int i = 0; i < ni; i++) {
    int64_t a = (int64_t)(c->iux[i] - c->ilx[i]) * (int64_t)(c->iuy[i] - c->ily[i]);
    if (i >= nr) {
      c->nonplace.push_back(i);
      c->nonplace_area += a;
    } else if (c->fixed[i]) {
      int rlx = std::max(c->core_lx, c->ilx[i]), rly = std::max(c->core_ly, c->ily[i]);
      int rux = std::min(c->core_ux, c->iux[i]), ruy = std::min(c->core_uy, c->iuy[i]);
      if (!(rlx >= rux || rly >= ruy)) {
        c->nonplace.push_back(i);
        c->nonplace_area += (int64_t)(rux - rlx) * (int64_t)(ruy - rly);
      }
    } else {
      c->inst2place[i] = (int)c->place.size();
      c->place.push_back(i);
      c->place_area += a;
      if (c->iuy[i] - c->ily[i] > c->site_h * 6)
        c->macro_area += a;
      else
        c->std_area += a;
    }
  }
  // pins (src/placerBase.cpp:330-406, 702-713)
  c->n_net = nl->num_nets;
  c->n_pin = nl->num_pins;
  const int np = c->n_pin;
  c->net_start.assign(nl->net_pin_start, nl->net_pin_start + c->n_net + 1);
  c->pin_inst.resize(np);
  c->pin_offx.assign(np, 0);
  c->pin_offy.assign(np, 0);
  c->pin_cx.resize(np);
  c->pin_cy.resize(np);
  c->pin_box.assign((size_t)4 * np, 0);
  for (int q = 0; q < np; q++) {
    int i = nl->pin_inst[q];
    if (i >= nr || i < -1) return fail(*out = c.release(), RP_ERR_ARG, "pin %d: bad instance %d", q, i);
    c->pin_inst[q] = i;
    if (i >= 0) {
      c->pin_offx[q] = nl->pin_off[2 * q];
      c->pin_offy[q] = nl->pin_off[2 * q + 1];
      c->pin_cx[q] = c->ilx[i] + c->mw[i] / 2 + c->pin_offx[q];
      c->pin_cy[q] = c->ily[i] + c->mh[i] / 2 + c->pin_offy[q];
    } else {
      int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
      if (nl->pin_box) {
        lx = nl->pin_box[4 * q];
        ly = nl->pin_box[4 * q + 1];
        ux = nl->pin_box[4 * q + 2];
        uy = nl->pin_box[4 * q + 3];
      }
      c->pin_box[4 * (size_t)q] = lx;
      c->pin_box[4 * (size_t)q + 1] = ly;
      c->pin_box[4 * (size_t)q + 2] = ux;
      c->pin_box[4 * (size_t)q + 3] = uy;
      c->pin_cx[q] = (int)(((int64_t)lx + ux) / 2);
      c->pin_cy[q] = (int)(((int64_t)ly + uy) / 2);
      // Pin::updateCoordi(dbBTerm*) warns and takes (0, 0) (src/placerBase.cpp:394-399)
      if (lx == INT_MAX || ly == INT_MAX || ux == INT_MIN || uy == INT_MIN) c->n_unplaced_ports++;
    }
  }
  // per-instance pin order (src/placerBase.cpp:728-743)
  c->ipin_start.assign(ni + 1, 0);
  if (nl->inst_pin_start && nl->inst_pins) {
    for (int i = 0; i <= nr; i++) c->ipin_start[i] = nl->inst_pin_start[i];
    for (int i = nr + 1; i <= ni; i++) c->ipin_start[i] = c->ipin_start[nr];
    c->ipins.assign(nl->inst_pins, nl->inst_pins + nl->inst_pin_start[nr]);
  } else {
    std::vector<int> cnt(ni + 1, 0);
    for (int q = 0; q < np; q++)
      if (c->pin_inst[q] >= 0) cnt[c->pin_inst[q] + 1]++;
    for (int i = 0; i < ni; i++) cnt[i + 1] += cnt[i];
    c->ipin_start = cnt;
    c->ipins.resize(cnt[ni]);
    for (int q = 0; q < np; q++)
      if (c->pin_inst[q] >= 0) c->ipins[cnt[c->pin_inst[q]]++] = q;
  }
  *out = c.release();
  // utilisation check of PlacerBase::printInfo (src/placerBase.cpp:936-956)
  rp_ctx* cc = *out;
  int64_t coreArea = (int64_t)(cc->core_ux - cc->core_lx) * (int64_t)(cc->core_uy - cc->core_ly);
  float util = (float)cc->place_area / (float)(coreArea - cc->nonplace_area) * 100;
  if (util >= 100.1) return fail(cc, RP_ERR_FILLER, "Util exceeds 100%%: %f", util);
  return RP_OK;
}

extern "C" const char* rp_last_error(rp_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int rp_get_pb_info(rp_ctx* c, rp_pb_info* o) {
  if (!c || !o) return RP_ERR_ARG;
  memset(o, 0, sizeof *o);
  o->dbu_per_micron = c->dbu;
  o->site_w = c->site_w;
  o->site_h = c->site_h;
  o->die_lx = c->die_lx;
  o->die_ly = c->die_ly;
  o->die_ux = c->die_ux;
  o->die_uy = c->die_uy;
  o->core_lx = c->core_lx;
  o->core_ly = c->core_ly;
  o->core_ux = c->core_ux;
  o->core_uy = c->core_uy;
  const int ni = (int)c->ilx.size();
  o->num_insts = ni;
  o->num_place_insts = (int)c->place.size();
  o->num_dummy_insts = ni - c->n_real;
  o->num_fixed_insts = (int)c->nonplace.size() - o->num_dummy_insts;
  o->num_nets = c->n_net;
  o->num_pins = c->n_pin;
  o->num_unplaced_ports = c->n_unplaced_ports;
  // PlacerBase::printInfo (src/placerBase.cpp:936-950)
  o->core_area = (int64_t)(c->core_ux - c->core_lx) * (int64_t)(c->core_uy - c->core_ly);
  o->nonplace_insts_area = c->nonplace_area;
  o->place_insts_area = c->place_area;
  o->std_insts_area = c->std_area;
  o->macro_insts_area = c->macro_area;
  o->util = (float)c->place_area / (float)(o->core_area - c->nonplace_area) * 100;
  return RP_OK;
}

extern "C" void rp_destroy(rp_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->h_state) rp_pinned_free(c->h_state);
  const hipStream_t st = c->stream;
  const int dev = c->device;
  delete c;  // its device buffers go back to the pool, ordered after the synchronised stream's work
  rp_stream_release(dev, st);
}


// ===========================================================================
// InitialPlace: host B2B assembly, device BiCGSTAB
// ===========================================================================
namespace {

struct Csr {
  int n = 0;
  std::vector<int> rowp, col;
  std::vector<float> val;
};

struct Trip {
  int r, c;
  float v;
};

// setFromTriplets: duplicates summed in triplet order (per row, by column).
// Stable counting sort by row, then a stable in-place insertion sort of each
// (short) row by column; equal (row, col) entries are added in triplet order.
void build_csr(int n, const std::vector<Trip>& t, Csr& m) {
  m.n = n;
  std::vector<int> start(n + 1, 0);
  for (const Trip& x : t) start[x.r + 1]++;
  for (int r = 0; r < n; r++) start[r + 1] += start[r];
  std::vector<int> tc(t.size());
  std::vector<float> tv(t.size());
  {
    std::vector<int> fillp(start.begin(), start.end() - 1);
    for (const Trip& x : t) {
      int at = fillp[x.r]++;
      tc[at] = x.c;
      tv[at] = x.v;
    }
  }
  m.rowp.assign(n + 1, 0);
  m.col.resize(t.size());
  m.val.resize(t.size());
  int nnz = 0;
  for (int r = 0; r < n; r++) {
    int a = start[r], b = start[r + 1];
    for (int k = a + 1; k < b; k++) {  // stable insertion sort by column
      int ck = tc[k];
      float vk = tv[k];
      int j = k - 1;
      while (j >= a && tc[j] > ck) {
        tc[j + 1] = tc[j];
        tv[j + 1] = tv[j];
        j--;
      }
      tc[j + 1] = ck;
      tv[j + 1] = vk;
    }
    for (int k = a; k < b;) {
      int cc = tc[k];
      float v = 0.0f;
      while (k < b && tc[k] == cc) v += tv[k++];
      m.col[nnz] = cc;
      m.val[nnz] = v;
      nnz++;
    }
    m.rowp[r + 1] = nnz;
  }
  m.col.resize(nnz);
  m.val.resize(nnz);
}

}  // namespace

// Locality order of the movable instances: breadth-first over the cell-net
// graph (cells reached through a shared net get neighbouring indices), roots
// in index order.  Device arrays of cells use it so that gathers through nets
// stay within cache lines.  Depends on the netlist only; computed once.
const std::vector<int>& rp_movable_order(rp_ctx* c) {
  const int n = (int)c->place.size(), nn = c->n_net, npn = c->n_pin;
  if ((int)c->bfs_order.size() == n) return c->bfs_order;
  std::vector<int> cell_net_start(n + 1, 0), cell_nets;
  for (int g = 0; g < n; g++) {
    int i = c->place[g];
    cell_net_start[g + 1] = cell_net_start[g] + (c->ipin_start[i + 1] - c->ipin_start[i]);
  }
  std::vector<int> pin_net_ref(npn);
  for (int e = 0; e < nn; e++)
    for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) pin_net_ref[q] = e;
  cell_nets.resize(cell_net_start[n]);
  for (int g = 0; g < n; g++) {
    int i = c->place[g], o = cell_net_start[g];
    for (int k = c->ipin_start[i]; k < c->ipin_start[i + 1]; k++) cell_nets[o++] = pin_net_ref[c->ipins[k]];
  }
  std::vector<uint8_t> cell_seen(n, 0), net_seen(nn, 0);
  std::vector<int>& order = c->bfs_order;
  order.clear();
  order.reserve(n);
  for (int root = 0; root < n; root++) {
    if (cell_seen[root]) continue;
    size_t head = order.size();
    order.push_back(root);
    cell_seen[root] = 1;
    while (head < order.size()) {
      int g = order[head++];
      for (int k = cell_net_start[g]; k < cell_net_start[g + 1]; k++) {
        int e = cell_nets[k];
        if (net_seen[e]) continue;
        net_seen[e] = 1;
        for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) {
          int i = c->pin_inst[q];
          int h = i >= 0 ? c->inst2place[i] : -1;
          if (h >= 0 && !cell_seen[h]) {
            cell_seen[h] = 1;
            order.push_back(h);
          }
        }
      }
    }
  }
  return order;
}

// Sequential host assembly of the B2B systems (updatePinInfo,
// src/initialPlace.cpp:128-182, and createSparseMatrix, :185-378) from the
// device's current pin and instance centres, in the device's row order
// (a.h_pin_k); uploads the CSR and right-hand sides into `a`.  The device assembly builds the same bits; this path exists
// so that tests can check it.
int rp_ip_host_assemble(rp_ctx* c, RpIpAsm& a) {
  const rp_params& p = c->prm;
  const int n = a.n, np = c->n_pin;
  hipStream_t st = c->stream;
  std::vector<int> pcx(np), pcy(np), icx(n), icy(n);
  HIPCHK(c, hipMemcpyAsync(pcx.data(), a.pin_cx.p, sizeof(int) * np, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(pcy.data(), a.pin_cy.p, sizeof(int) * np, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(icx.data(), a.icx.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(icy.data(), a.icy.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  // updatePinInfo: the first pin reaching a strict new extreme keeps the flag
  std::vector<uint8_t> mm(np, 0);
  for (int e = 0; e < c->n_net; e++) {
    int pminx = -1, pmaxx = -1, pminy = -1, pmaxy = -1;
    int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
    for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) {
      if (lx > pcx[q]) {
        if (pminx >= 0) mm[pminx] &= (uint8_t)~1u;
        lx = pcx[q];
        pminx = q;
        mm[q] |= 1u;
      }
      if (ux < pcx[q]) {
        if (pmaxx >= 0) mm[pmaxx] &= (uint8_t)~2u;
        ux = pcx[q];
        pmaxx = q;
        mm[q] |= 2u;
      }
      if (ly > pcy[q]) {
        if (pminy >= 0) mm[pminy] &= (uint8_t)~4u;
        ly = pcy[q];
        pminy = q;
        mm[q] |= 4u;
      }
      if (uy < pcy[q]) {
        if (pmaxy >= 0) mm[pmaxy] &= (uint8_t)~8u;
        uy = pcy[q];
        pmaxy = q;
        mm[q] |= 8u;
      }
    }
  }
  // createSparseMatrix
  std::vector<float> bx(n, 0.0f), by(n, 0.0f);
  std::vector<Trip> tx, ty;
  const float minDiff = (float)p.initial_place_min_diff_length;
  auto pair = [&](int p1, int p2, int k1, int k2, const std::vector<int>& ic, float netWeight,
                  std::vector<Trip>& t, std::vector<float>& b) {
    int diff = std::abs(p1 - p2);
    float w = diff > p.initial_place_min_diff_length ? netWeight / (float)diff : netWeight / minDiff;
    if (k1 >= 0 && k2 >= 0) {
      t.push_back({k1, k1, w});
      t.push_back({k2, k2, w});
      t.push_back({k1, k2, -w});
      t.push_back({k2, k1, -w});
      b[k1] += -w * (float)((p1 - ic[k1]) - (p2 - ic[k2]));
      b[k2] += -w * (float)((p2 - ic[k2]) - (p1 - ic[k1]));
    } else if (k1 < 0 && k2 >= 0) {
      t.push_back({k2, k2, w});
      b[k2] += w * (float)(p1 - (p2 - ic[k2]));
    } else if (k1 >= 0 && k2 < 0) {
      t.push_back({k1, k1, w});
      b[k1] += w * (float)(p2 - (p1 - ic[k1]));
    }
  };
  for (int e = 0; e < c->n_net; e++) {
    int s0 = c->net_start[e], deg = c->net_start[e + 1] - s0;
    if (deg <= 1 || deg >= p.initial_place_max_fanout) continue;
    float netWeight = p.initial_place_net_weight_scale / (float)(deg - 1);
    for (int a1 = 0; a1 < deg; a1++) {
      int q1 = s0 + a1, i1 = c->pin_inst[q1];
      int k1 = a.h_pin_k[q1];
      for (int a2 = 0; a2 <= a1; a2++) {
        int q2 = s0 + a2, i2 = c->pin_inst[q2];
        if (i1 == i2) continue;
        int k2 = a.h_pin_k[q2];
        if ((mm[q1] & 3u) || (mm[q2] & 3u)) pair(pcx[q1], pcx[q2], k1, k2, icx, netWeight, tx, bx);
        if ((mm[q1] & 12u) || (mm[q2] & 12u)) pair(pcy[q1], pcy[q2], k1, k2, icy, netWeight, ty, by);
      }
    }
  }
  Csr M[2];
  build_csr(n, tx, M[0]);
  build_csr(n, ty, M[1]);
  const std::vector<float>* B[2] = {&bx, &by};
  for (int ax = 0; ax < 2; ax++) {
    HIPCHK(c, hipMemcpyAsync(a.rowp[ax].p, M[ax].rowp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, st));
    HIPCHK(c, dupload(a.col[ax], M[ax].col, st));
    HIPCHK(c, dupload(a.val[ax], M[ax].val, st));
    HIPCHK(c, hipMemcpyAsync(a.b[ax].p, B[ax]->data(), sizeof(float) * n, hipMemcpyHostToDevice, st));
  }
  HIPCHK(c, hipStreamSynchronize(st));
  return 0;
}

// The steps of rp_initial_place around its iterations, for a driver that
// runs the iterations of many contexts together (rp_batch.cpp): *run is the
// context's run, or null when there is nothing to iterate.
static void ip_prepare(rp_ctx* c) {
  const int n = (int)c->place.size();
  c->ip_err.clear();
  c->ip_hpwl.clear();
  // placeInstsCenter (src/initialPlace.cpp:108-115)
  if (!c->prm.incremental_place_mode) {
    int ccx = (c->core_lx + c->core_ux) / 2, ccy = (c->core_ly + c->core_uy) / 2;
    for (int k = 0; k < n; k++) inst_set_center(c, c->place[k], ccx, ccy);
  }
}

int rp_initial_place_begin(rp_ctx* c, hipStream_t run_stream, RpIpRun** run) {
  *run = nullptr;
  ip_prepare(c);
  if (c->prm.initial_place_max_iter > 0 && !c->place.empty()) return rp_ipr_begin(c, run_stream, false, run);
  return 0;
}

int rp_initial_place_end(rp_ctx* c, RpIpRun* run) {
  const int r = run ? rp_ipr_end(run) : 0;
  c->nb_ready = false;
  return r;
}

extern "C" int rp_initial_place(rp_ctx* c, rp_ip_stats* stats) {
  if (!c) return RP_ERR_ARG;
  auto t0 = std::chrono::steady_clock::now();
  const rp_params& p = c->prm;
  const int n = (int)c->place.size();
  ip_prepare(c);
  if (p.initial_place_max_iter > 0 && n > 0) {
    // B2B assembly on the device; RP_IP_HOST_ASSEMBLY=1 selects the sequential
    // host assembly of the same systems
    const bool host_asm = getenv("RP_IP_HOST_ASSEMBLY") != nullptr;
    double t_asm = 0, t_solve = 0;
    int r = rp_ip_device_loop(c, host_asm, &t_asm, &t_solve);
    if (getenv("RP_IP_TIMING"))
      fprintf(stderr, "[rp] initial place phases (%s assembly): B2B assembly %.3fs, BiCGSTAB + update %.3fs\n",
              host_asm ? "host" : "device", t_asm, t_solve);
    if (r) return r;
  }
  c->nb_ready = false;
  if (stats) {
    stats->iterations = (int)c->ip_err.size();
    stats->last_error = c->ip_err.empty() ? 0.0f : c->ip_err.back();
    stats->last_hpwl = c->ip_hpwl.empty() ? 0 : c->ip_hpwl.back();
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return RP_OK;
}

extern "C" int rp_get_ip_records(rp_ctx* c, float* err, int64_t* hpwl, int32_t cap, int32_t* count) {
  if (!c || !count) return RP_ERR_ARG;
  *count = (int32_t)c->ip_err.size();
  for (int k = 0; k < std::min<int>(cap, *count); k++) {
    if (err) err[k] = c->ip_err[k];
    if (hpwl) hpwl[k] = c->ip_hpwl[k];
  }
  return RP_OK;
}

// ===========================================================================
// NesterovBase::init and device upload
// ===========================================================================
static void make_twiddles(int n, std::vector<float2>& fft, std::vector<float2>& dct) {
  fft.resize(std::max(1, n / 2));
  dct.resize(n);
  for (int k = 0; k < n / 2; k++) {
    double a = -2.0 * M_PI * (double)k / (double)n;
    fft[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  for (int k = 0; k < n; k++) {
    double a = M_PI * (double)k / (2.0 * (double)n);
    dct[k] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
}

static bool is_pow2(int v) { return v >= 4 && v <= 2048 && (v & (v - 1)) == 0; }

extern "C" int rp_nesterov_setup(rp_ctx* c, rp_nb_info* info) {
  if (!c) return RP_ERR_ARG;
  const rp_params& p = c->prm;
  const int ngi = (int)c->place.size();
  if (ngi == 0) return fail(c, RP_ERR_ARG, "no movable instances");
  // fillers (initFillerGCells, src/nesterovBase.cpp:1066-1155)
  std::vector<int> dxs(ngi), dys(ngi);
  for (int k = 0; k < ngi; k++) {
    int i = c->place[k];
    dxs[k] = c->iux[i] - c->ilx[i];
    dys[k] = c->iuy[i] - c->ily[i];
  }
  std::sort(dxs.begin(), dxs.end());
  std::sort(dys.begin(), dys.end());
  int minIdx = (int)((double)ngi * 0.05), maxIdx = (int)((double)ngi * 0.95);
  int64_t dxSum = 0, dySum = 0;
  for (int k = minIdx; k < maxIdx; k++) {
    dxSum += dxs[k];
    dySum += dys[k];
  }
  int avgDx = maxIdx > minIdx ? (int)(dxSum / (maxIdx - minIdx)) : 0;
  int avgDy = maxIdx > minIdx ? (int)(dySum / (maxIdx - minIdx)) : 0;
  int cdx = c->core_ux - c->core_lx, cdy = c->core_uy - c->core_ly;
  int64_t coreArea = (int64_t)cdx * (int64_t)cdy;
  int64_t white = coreArea - c->nonplace_area;
  int64_t movable = (int64_t)((float)white * p.target_density);
  int64_t totalFiller = movable - c->std_area - (int64_t)((float)c->macro_area * p.target_density);
  if (totalFiller < 0)
    return fail(c, RP_ERR_FILLER,
                "Filler area is negative!!\n Please put higher target density or \n Re-floorplan to have enough "
                "coreArea\n");
  int fillerCnt = (avgDx * avgDy) > 0 ? (int)(totalFiller / (int64_t)(avgDx * avgDy)) : 0;
  c->nb_core_area = coreArea;
  c->nb_white = white;
  c->nb_movable = movable;
  c->nb_total_filler = totalFiller;
  c->filler_dx = avgDx;
  c->filler_dy = avgDy;
  c->n_gi = ngi;
  c->n_gf = fillerCnt;
  c->n_g = ngi + fillerCnt;
  const int ng = c->n_g;
  std::vector<int> glx(ng), gly(ng), gux(ng), guy(ng);
  c->g_inst.assign(ng, -1);
  c->g_kind.assign(ng, 2);
  for (int k = 0; k < ngi; k++) {
    int i = c->place[k];
    c->g_inst[k] = i;
    glx[k] = c->ilx[i];
    gly[k] = c->ily[i];
    gux[k] = c->iux[i];
    guy[k] = c->iuy[i];
    c->g_kind[k] = (c->iuy[i] - c->ily[i]) > c->site_h * 6 ? 1 : 0;
  }
  std::mt19937 rv(0);
  for (int k = 0; k < fillerCnt; k++) {
    auto rx = rv();
    auto ry = rv();
    int cx = (int)(rx % (unsigned long)(long)cdx + (unsigned long)(long)c->core_lx);
    int cy = (int)(ry % (unsigned long)(long)cdy + (unsigned long)(long)c->core_ly);
    int g = ngi + k;
    glx[g] = cx - avgDx / 2;
    gly[g] = cy - avgDy / 2;
    gux[g] = cx + avgDx / 2;
    guy[g] = cy + avgDy / 2;
  }
  // bin grid (BinGrid::initBins, src/nesterovBase.cpp:631-717)
  c->td = p.target_density;
  int64_t totalBinArea = (int64_t)cdx * (int64_t)cdy;
  int64_t avgArea = c->place_area / (int64_t)ngi;
  int64_t idealBinArea = (int64_t)std::round((float)avgArea / c->td);
  int idealBinCnt = idealBinArea > 0 ? (int)(totalBinArea / idealBinArea) : 0;
  c->nb_avg_area = avgArea;
  c->nb_ideal_bin_area = idealBinArea;
  c->nb_ideal_bin_cnt = idealBinCnt;
  int found = 2;
  for (found = 2; found <= 1024; found *= 2)
    if (found * found <= idealBinCnt && 4 * found * found > idealBinCnt) break;
  c->nbx = p.bin_grid_cnt_x ? p.bin_grid_cnt_x : found;
  c->nby = p.bin_grid_cnt_y ? p.bin_grid_cnt_y : found;
  if (!is_pow2(c->nbx) || !is_pow2(c->nby))
    return fail(c, RP_ERR_ARG, "bin grid %dx%d: each side must be a power of two in [4, 2048]", c->nbx, c->nby);
  c->bsx = (int)std::ceil((float)cdx / (float)c->nbx);
  c->bsy = (int)std::ceil((float)cdy / (float)c->nby);
  // the kernels assume the regular layout the reference's loop produces
  if ((int64_t)(c->nbx - 1) * c->bsx >= cdx || (int64_t)(c->nby - 1) * c->bsy >= cdy)
    return fail(c, RP_ERR_ARG, "bin grid %dx%d does not tile the core regularly", c->nbx, c->nby);
  const int nb = c->nbx * c->nby;
  // nonPlace areas (updateBinsNonPlaceArea, src/nesterovBase.cpp:719-741)
  c->nonplace_bins.assign(nb, 0);
  for (int i : c->nonplace) {
    auto px = site_range(c->ilx[i], c->iux[i], c->core_lx, c->bsx, 0, c->nbx);
    auto py = site_range(c->ily[i], c->iuy[i], c->core_ly, c->bsy, 0, c->nby);
    for (int a = px.first; a < px.second; a++) {
      int blx = c->core_lx + a * c->bsx, bux = std::min(blx + c->bsx, c->core_ux);
      for (int b = py.first; b < py.second; b++) {
        int bly = c->core_ly + b * c->bsy, buy = std::min(bly + c->bsy, c->core_uy);
        int rlx = std::max(blx, c->ilx[i]), rly = std::max(bly, c->ily[i]);
        int rux = std::min(bux, c->iux[i]), ruy = std::min(buy, c->iuy[i]);
        int64_t ov = (rlx >= rux || rly >= ruy) ? 0 : (int64_t)(rux - rlx) * (int64_t)(ruy - rly);
        c->nonplace_bins[(size_t)b * c->nbx + a] += (int64_t)((float)ov * c->td);
      }
    }
  }
  // density size and scale (src/nesterovBase.cpp:1033-1061), then the
  // layout clamp of NesterovPlace::init (src/nesterovBase.cpp:1245-1268)
  c->g_dx.resize(ng);
  c->g_dy.resize(ng);
  c->g_dhx.resize(ng);
  c->g_dhy.resize(ng);
  c->g_scale.resize(ng);
  c->g_icx.resize(ng);
  c->g_icy.resize(ng);
  for (int g = 0; g < ng; g++) {
    int gdx = gux[g] - glx[g], gdy = guy[g] - gly[g];
    c->g_dx[g] = gdx;
    c->g_dy[g] = gdy;
    float scaleX, scaleY, dsx, dsy;
    if ((long double)gdx < kSqrt2 * (long double)c->bsx) {
      scaleX = (float)gdx / (float)(kSqrt2 * (long double)c->bsx);
      dsx = (float)(kSqrt2 * (long double)(float)c->bsx);
    } else {
      scaleX = 1.0f;
      dsx = (float)gdx;
    }
    if ((long double)gdy < kSqrt2 * (long double)c->bsy) {
      scaleY = (float)gdy / (float)(kSqrt2 * (long double)c->bsy);
      dsy = (float)(kSqrt2 * (long double)(float)c->bsy);
    } else {
      scaleY = 1.0f;
      dsy = (float)gdy;
    }
    int ddx = (int)dsx, ddy = (int)dsy;
    int ccx = (gux[g] + glx[g]) / 2, ccy = (guy[g] + gly[g]) / 2;
    int dlx = ccx - ddx / 2, dly = ccy - ddy / 2;
    int dux = ccx + ddx / 2, duy = ccy + ddy / 2;
    c->g_scale[g] = scaleX * scaleY;
    int wdx = dux - dlx, wdy = duy - dly;
    float tlx = (float)dlx, tly = (float)dly;
    if (tlx < (float)c->core_lx) tlx = (float)c->core_lx;
    if (tly < (float)c->core_ly) tly = (float)c->core_ly;
    if (tlx + (float)wdx > (float)c->core_ux) tlx = (float)(c->core_ux - wdx);
    if (tly + (float)wdy > (float)c->core_uy) tly = (float)(c->core_uy - wdy);
    int nlx = (int)tlx, nly = (int)tly;
    c->g_dhx[g] = wdx / 2;
    c->g_dhy[g] = wdy / 2;
    c->g_icx[g] = (float)((nlx + nlx + wdx) / 2);
    c->g_icy[g] = (float)((nly + nly + wdy) / 2);
  }

  // ---------------- device upload ----------------
  hipStream_t s = c->stream;
  // ---- locality order of the device arrays ----
  // Movable cells are numbered in breadth-first order over the netlist
  // (cells that share nets get nearby indices, so the scatter's per-block
  // aggregation and the pin gathers of the WA / gradient kernels touch
  // nearby memory); fillers keep their order after them.  Nets are grouped
  // by degree bucket (equal work per wavefront) and ordered by the lowest
  // new index of their cells.  Pin order inside a net and inside a cell is
  // the reference order, so every per-net and per-cell float sum keeps its
  // reference summation order.
  const int nn = c->n_net, npn = c->n_pin;
  {
    const std::vector<int>& order = rp_movable_order(c);
    c->gperm.assign(order.begin(), order.end());
    for (int g = ngi; g < ng; g++) c->gperm.push_back(g);
    c->ginv.assign(ng, 0);
    for (int k = 0; k < ng; k++) c->ginv[c->gperm[k]] = k;
  }
  std::vector<int2> dhalf(ng);
  std::vector<float> area(ng), npin(ng), scale(ng);
  std::vector<uint8_t> kind(ng);
  for (int k = 0; k < ng; k++) {
    int g = c->gperm[k];
    dhalf[k] = make_int2(c->g_dhx[g], c->g_dhy[g]);
    area[k] = (float)c->g_dx[g] * (float)c->g_dy[g];
    int i = c->g_inst[g];
    npin[k] = i >= 0 ? (float)(c->ipin_start[i + 1] - c->ipin_start[i]) : 0.0f;
    scale[k] = c->g_scale[g];
    kind[k] = c->g_kind[g];
  }
  auto bucket = [](int deg) {
    int b = 0;
    while (b < 12 && (2 << b) < deg) b++;  // 1-2, 3-4, 5-8, 9-16, ...
    return b;
  };
  std::vector<int> net_key(nn, INT_MAX);
  for (int e = 0; e < nn; e++)
    for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) {
      int i = c->pin_inst[q];
      int g = i >= 0 ? c->inst2place[i] : -1;
      if (g >= 0) net_key[e] = std::min(net_key[e], c->ginv[g]);
    }
  // nets in locality order, cut into one contiguous chunk per XCD (matching
  // the kernel's XCD-contiguous block ranges, see xcd_block), each chunk
  // sorted by degree bucket so every XCD gets the same mix of net sizes and
  // every wavefront nets of similar size
  c->net_perm.resize(nn);
  std::iota(c->net_perm.begin(), c->net_perm.end(), 0);
  std::stable_sort(c->net_perm.begin(), c->net_perm.end(),
                   [&](int a, int b) { return net_key[a] < net_key[b]; });
  {
    const int nblk = (nn + RP_BLOCK - 1) / RP_BLOCK, kXcd = 8;
    int per = nblk / kXcd, rem = nblk % kXcd, at = 0;
    for (int x = 0; x < kXcd && at < nn; x++) {
      int blocks = per + (x < rem ? 1 : 0);
      int end = std::min(nn, at + blocks * RP_BLOCK);
      std::stable_sort(c->net_perm.begin() + at, c->net_perm.begin() + end, [&](int a, int b) {
        return bucket(c->net_start[a + 1] - c->net_start[a]) < bucket(c->net_start[b + 1] - c->net_start[b]);
      });
      at = end;
    }
  }
  std::vector<int> nstart(nn + 1, 0);
  for (int k = 0; k < nn; k++) {
    int e = c->net_perm[k];
    nstart[k + 1] = nstart[k] + (c->net_start[e + 1] - c->net_start[e]);
  }
  std::vector<int> pcell(npn);
  std::vector<int2> ppos(npn);
  std::vector<int> pin_dev(npn);
  for (int k = 0; k < nn; k++) {
    int e = c->net_perm[k];
    for (int q = c->net_start[e], m = nstart[k]; q < c->net_start[e + 1]; q++, m++) {
      pin_dev[q] = m;
      int i = c->pin_inst[q];
      int g = i >= 0 ? c->inst2place[i] : -1;
      pcell[m] = g >= 0 ? c->ginv[g] : -1;
      ppos[m] = g >= 0 ? make_int2(c->pin_offx[q], c->pin_offy[q]) : make_int2(c->pin_cx[q], c->pin_cy[q]);
    }
  }
  std::vector<int> cp_start(ng + 1, 0), cp_pin;
  for (int k = 0; k < ng; k++) {
    int i = c->g_inst[c->gperm[k]];
    if (i >= 0)
      for (int j = c->ipin_start[i]; j < c->ipin_start[i + 1]; j++) {
        int q = c->ipins[j];
        cp_pin.push_back(pin_dev[q]);
      }
    cp_start[k + 1] = (int)cp_pin.size();
  }
  std::vector<float2> twx, tw2x, twy, tw2y;
  make_twiddles(c->nbx, twx, tw2x);
  make_twiddles(c->nby, twy, tw2y);
  // FFT::init twiddles with the reference's long double rounding (src/fft.cpp:78-90)
  std::vector<float> wx(c->nbx), wx2(c->nbx), wy(c->nby), wy2(c->nby);
  for (int i = 0; i < c->nbx; i++) {
    wx[i] = (float)(kPi * (long double)(float)i / (long double)(float)c->nbx);
    wx2[i] = wx[i] * wx[i];
  }
  for (int j = 0; j < c->nby; j++) {
    wy[j] = (float)(kPi * (long double)(float)j / (long double)(float)c->nby * (long double)(float)c->bsy /
                    (long double)(float)c->bsx);
    wy2[j] = wy[j] * wy[j];
  }
  HIPCHK(c, dupload(c->d_dhalf, dhalf, s));
  HIPCHK(c, dupload(c->d_scale, scale, s));
  HIPCHK(c, dupload(c->d_kind, kind, s));
  HIPCHK(c, dupload(c->d_area, area, s));
  HIPCHK(c, dupload(c->d_npin, npin, s));
  HIPCHK(c, dalloc(c->d_slp, (size_t)3 * ng));
  HIPCHK(c, dalloc(c->d_sumg, (size_t)3 * ng));
  HIPCHK(c, dalloc(c->d_co, (size_t)2 * ng));
  HIPCHK(c, dalloc(c->d_ic, ng));
  {
    std::vector<int> iota(ng);
    for (int k = 0; k < ng; k++) iota[k] = k;
    HIPCHK(c, dupload(c->d_sord, iota, s));  // any permutation is valid until the first refresh
    HIPCHK(c, dalloc(c->d_srank, std::max(1, ng)));
    HIPCHK(c, dalloc(c->d_scnt, RP_SORT_TILES));
    HIPCHK(c, dalloc(c->d_soff, RP_SORT_TILES));
    HIPCHK(c, hipMemsetAsync(c->d_scnt.p, 0, sizeof(int) * RP_SORT_TILES, s));
    RpSortBufs& sb = c->sortb;
    sb.cnt = c->d_scnt.p;
    sb.off = c->d_soff.p;
    sb.rank = c->d_srank.p;
    sb.sord = c->d_sord.p;
    sb.tile_shift = rp_sort_tile_shift(c->nbx, c->nby);
  }
  HIPCHK(c, hipMemsetAsync(c->d_sumg.p, 0, sizeof(float2) * 3 * ng, s));
  HIPCHK(c, dupload(c->d_cp_start, cp_start, s));
  HIPCHK(c, dupload(c->d_cp_pin, cp_pin, s));
  HIPCHK(c, dalloc(c->d_pgrad, std::max(1, npn)));
  HIPCHK(c, dupload(c->d_nstart, nstart, s));
  // wirelength workgroups: consecutive nets, at most RP_WA_NETS nets and
  // `cap` pins each; a net with more pins gets a workgroup of its own.  One
  // packing per tile-shape variant (256 / 512 / 1024 pins).
  int nwa_v[RP_WA_VARIANTS];
  for (int v = 0; v < RP_WA_VARIANTS; v++) {
    const int cap = 256 << v;
    std::vector<int4> wa_rng;  // nets [x, y), their pins [z, w)
    for (int e = 0; e < nn;) {
      const int e0 = e;
      if (nstart[e + 1] - nstart[e] > cap) {
        wa_rng.push_back(make_int4(e, e + 1, nstart[e], nstart[e + 1]));
        e++;
        continue;
      }
      while (e < nn && e - e0 < RP_WA_NETS && nstart[e + 1] - nstart[e0] <= cap) e++;
      wa_rng.push_back(make_int4(e0, e, nstart[e0], nstart[e]));
    }
    if (wa_rng.empty()) wa_rng.push_back(make_int4(0, 0, 0, 0));
    nwa_v[v] = (int)wa_rng.size();
    HIPCHK(c, dupload(c->d_wa_rng[v], wa_rng, s));
  }
  const int nwa = nwa_v[RP_WA_DEFAULT_VARIANT];
  const int nwa_max = *std::max_element(nwa_v, nwa_v + RP_WA_VARIANTS);
  HIPCHK(c, dupload(c->d_pcell, pcell, s));
  HIPCHK(c, dupload(c->d_ppos, ppos, s));
  HIPCHK(c, dalloc(c->d_nbox, nn));
  HIPCHK(c, dalloc(c->d_nsum, (size_t)2 * nn));
  HIPCHK(c, dalloc(c->d_bin_inst, nb));
  HIPCHK(c, dalloc(c->d_bin_filler, nb));
  HIPCHK(c, hipMemsetAsync(c->d_bin_inst.p, 0, sizeof(unsigned long long) * nb, s));
  HIPCHK(c, hipMemsetAsync(c->d_bin_filler.p, 0, sizeof(unsigned long long) * nb, s));
  std::vector<float> npf(nb);
  for (int b = 0; b < nb; b++) npf[b] = (float)c->nonplace_bins[b];  // int64 -> float as the density formulas convert it
  HIPCHK(c, dupload(c->d_bin_npf, npf, s));
  HIPCHK(c, dalloc(c->d_rho, nb));
  HIPCHK(c, dalloc(c->d_phi, nb));
  HIPCHK(c, dalloc(c->d_field, nb));
  HIPCHK(c, dalloc(c->d_t, nb));
  HIPCHK(c, dalloc(c->d_y1, nb));
  HIPCHK(c, dalloc(c->d_y2, nb));
  HIPCHK(c, dupload(c->d_twx, twx, s));
  HIPCHK(c, dupload(c->d_tw2x, tw2x, s));
  HIPCHK(c, dupload(c->d_twy, twy, s));
  HIPCHK(c, dupload(c->d_tw2y, tw2y, s));
  HIPCHK(c, dupload(c->d_wx, wx, s));
  HIPCHK(c, dupload(c->d_wx2, wx2, s));
  HIPCHK(c, dupload(c->d_wy, wy, s));
  HIPCHK(c, dupload(c->d_wy2, wy2, s));
  int npg = (ng + RP_BLOCK - 1) / RP_BLOCK, npb = c->nby / 2, npn2 = nwa_max;
  HIPCHK(c, dalloc(c->d_part_grad, (size_t)RP_P_COUNT * npg));
  HIPCHK(c, dalloc(c->d_part_bin, npb));
  HIPCHK(c, dalloc(c->d_part_net, std::max(1, npn2)));
  HIPCHK(c, hipMemsetAsync(c->d_part_net.p, 0, sizeof(long long) * std::max(1, npn2), s));
  HIPCHK(c, dalloc(c->d_state, 1));
  HIPCHK(c, dalloc(c->d_rec, std::max(1, p.nesterov_place_max_iter)));
  HIPCHK(c, dalloc(c->d_dbg_wl, ng));
  HIPCHK(c, dalloc(c->d_dbg_dens, ng));
  // verbose-3 trace buffers: a record per evaluation (init + iterations x trials)
  c->trace = p.verbose >= 3;
  if (c->trace) {
    const size_t cap = (size_t)std::max(1, p.nesterov_place_max_iter) * std::max(1, p.max_back_track) + 2;
    HIPCHK(c, dalloc(c->d_evrec, cap));
    HIPCHK(c, dalloc(c->d_part_gn, npg));
    HIPCHK(c, dalloc(c->d_part_phi, std::max(1, c->nby / 2)));
    HIPCHK(c, dalloc(c->d_dbg_area, nb));
    HIPCHK(c, hipMemsetAsync(c->d_part_phi.p, 0, sizeof(double) * std::max(1, c->nby / 2), s));
    HIPCHK(c, hipMemsetAsync(c->d_part_gn.p, 0, sizeof(double) * npg, s));
  }
  HIPCHK(c, hipStreamSynchronize(s));

  RpDev& d = c->dev;
  memset(&d, 0, sizeof d);
  d.nc = ng;
  d.nn = nn;
  d.np = npn;
  d.nbx = c->nbx;
  d.nby = c->nby;
  d.bsx = c->bsx;
  d.bsy = c->bsy;
  d.core_lx = c->core_lx;
  d.core_ly = c->core_ly;
  d.core_ux = c->core_ux;
  d.core_uy = c->core_uy;
  d.td = c->td;
  d.scatter_per = ng >= RP_SS_WIDE_CELLS ? 2 : 1;
  d.fuse_scatter_wa = getenv("RP_NO_FUSE") == nullptr ? 1 : 0;  // RP_NO_FUSE=1: two launches
  if (const char* e = getenv("RP_SCATTER_PER")) {
    const int v = atoi(e);
    if (v == 1 || v == 2) d.scatter_per = v;
  }
  d.bar = p.min_wire_length_force_bar;
  d.min_precond = p.min_preconditioner;
  d.dhalf = c->d_dhalf.p;
  d.scale = c->d_scale.p;
  d.kind = c->d_kind.p;
  d.area = c->d_area.p;
  d.npin = c->d_npin.p;
  d.slp = c->d_slp.p;
  d.sumg = c->d_sumg.p;
  d.co = c->d_co.p;
  d.ic = c->d_ic.p;
  d.sord = c->d_sord.p;
  d.cp_start = c->d_cp_start.p;
  d.cp_pin = c->d_cp_pin.p;
  d.pgrad = c->d_pgrad.p;
  d.nstart = c->d_nstart.p;
  for (int v = 0; v < RP_WA_VARIANTS; v++) {
    d.nwa_v[v] = nwa_v[v];
    d.wa_rng_v[v] = c->d_wa_rng[v].p;
  }
  d.nwa = nwa;
  d.wa_rng = c->d_wa_rng[RP_WA_DEFAULT_VARIANT].p;
  d.pcell = c->d_pcell.p;
  d.ppos = c->d_ppos.p;
  d.nbox = c->d_nbox.p;
  d.nsum = c->d_nsum.p;
  d.bin_inst = c->d_bin_inst.p;
  d.bin_filler = c->d_bin_filler.p;
  d.bin_npf = c->d_bin_npf.p;
  d.store_rho = p.poisson_transform == 1 ? 1 : 0;
  d.rho = c->d_rho.p;
  d.field = c->d_field.p;
  d.phi = c->d_phi.p;
  d.t = c->d_t.p;
  d.y1 = c->d_y1.p;
  d.y2 = c->d_y2.p;
  d.twx = c->d_twx.p;
  d.tw2x = c->d_tw2x.p;
  d.twy = c->d_twy.p;
  d.tw2y = c->d_tw2y.p;
  d.wx = c->d_wx.p;
  d.wx2 = c->d_wx2.p;
  d.wy = c->d_wy.p;
  d.wy2 = c->d_wy2.p;
  d.part_grad = c->d_part_grad.p;
  d.part_bin = c->d_part_bin.p;
  d.npart_grad = npg;
  d.npart_bin = npb;
  d.part_net = c->d_part_net.p;
  d.st = c->d_state.p;
  d.rec = c->d_rec.p;
  d.dbg_wl = nullptr;
  d.dbg_dens = nullptr;
  if (c->trace) {
    d.evrec = c->d_evrec.p;
    d.evrec_cap = (int)c->d_evrec.n;
    d.part_gn = c->d_part_gn.p;
    d.part_phi = p.poisson_transform == 1 ? nullptr : c->d_part_phi.p;
    d.dbg_area = c->d_dbg_area.p;
  }

  RpState& S = *c->h_state;
  memset(&S, 0, sizeof S);
  S.num_cells = ng;
  S.max_iter = p.nesterov_place_max_iter;
  S.max_back_track = p.max_back_track;
  S.num_partials_grad = npg;
  S.num_partials_bin = npb;
  S.num_partials_net = nwa;
  S.target_overflow = p.target_overflow;
  S.min_phi_coef = p.min_phi_coef;
  S.reference_hpwl = p.reference_hpwl;
  S.init_density_penalty = p.init_density_penalty;
  S.init_wl_coef = p.init_wire_length_coef;
  S.area_denom = (float)c->std_area + (float)c->macro_area * c->td;  // int64 + float -> float
  S.bin_size_sum = (float)(c->bsx + c->bsy);
  S.initial_prev_coef = p.initial_prev_coordi_update_coef;
  S.max_phi_coef = p.max_phi_coef;
  S.ring.cur = 0;
  S.ring.next = 1;
  S.ring.prev = 2;
  S.ring.co_cur = 0;
  S.ring.co_next = 1;
  // Poisson transform: FP32 FFT (default) or split-BF16 GEMMs on matrix cores
  c->mfma.ready = false;
  c->dev.mfma = 0;
  if (p.poisson_transform == 1) {
    if (int r = rp_mfma_setup(c)) return r;
    c->mfma_view = c->mfma.view();
    c->dev.mfma = 1;
  } else if (p.poisson_transform != 0) {
    return fail(c, RP_ERR_ARG, "poisson_transform must be 0 (FFT) or 1 (matrix cores)");
  }
  c->nb_ready = true;
  if (info) {
    memset(info, 0, sizeof *info);
    info->num_gcells = ng;
    info->num_fillers = fillerCnt;
    info->num_place_insts = ngi;
    info->num_gnets = nn;
    info->num_gpins = npn;
    info->bin_cnt_x = c->nbx;
    info->bin_cnt_y = c->nby;
    info->bin_size_x = c->bsx;
    info->bin_size_y = c->bsy;
    info->core_lx = c->core_lx;
    info->core_ly = c->core_ly;
    info->core_ux = c->core_ux;
    info->core_uy = c->core_uy;
    info->filler_dx = avgDx;
    info->filler_dy = avgDy;
    info->std_insts_area = c->std_area;
    info->macro_insts_area = c->macro_area;
    info->place_insts_area = c->place_area;
    info->nonplace_insts_area = c->nonplace_area;
    info->target_density = c->td;
    info->core_area = c->nb_core_area;
    info->white_space_area = c->nb_white;
    info->movable_area = c->nb_movable;
    info->total_filler_area = c->nb_total_filler;
    info->average_place_inst_area = c->nb_avg_area;
    info->ideal_bin_area = c->nb_ideal_bin_area;
    info->ideal_bin_count = c->nb_ideal_bin_cnt;
    info->total_bin_area = totalBinArea;
  }
  return RP_OK;
}

extern "C" int rp_get_gcells(rp_ctx* c, int32_t* dhx, int32_t* dhy, float* scale, int32_t* dx, int32_t* dy,
                             float* icx, float* icy, uint8_t* kind) {
  if (!c || !c->nb_ready) return RP_ERR_STATE;
  for (int g = 0; g < c->n_g; g++) {
    if (dhx) dhx[g] = c->g_dhx[g];
    if (dhy) dhy[g] = c->g_dhy[g];
    if (scale) scale[g] = c->g_scale[g];
    if (dx) dx[g] = c->g_dx[g];
    if (dy) dy[g] = c->g_dy[g];
    if (icx) icx[g] = c->g_icx[g];
    if (icy) icy[g] = c->g_icy[g];
    if (kind) kind[g] = c->g_kind[g];
  }
  return RP_OK;
}

extern "C" int rp_get_nonplace(rp_ctx* c, int64_t* out) {
  if (!c || !c->nb_ready || !out) return RP_ERR_STATE;
  memcpy(out, c->nonplace_bins.data(), sizeof(int64_t) * c->nonplace_bins.size());
  return RP_OK;
}

// ===========================================================================
// stage-level evaluation (tests) and Poisson-only entry
// ===========================================================================
extern "C" int rp_eval_gradient(rp_ctx* c, const float* cx, const float* cy, float lam, float wlx, float wly,
                                rp_eval_out* o) {
  if (!c || !c->nb_ready || !cx || !cy || !o) return RP_ERR_STATE;
  const int ng = c->n_g, nb = c->nbx * c->nby, nn = c->n_net;
  hipStream_t s = c->stream;
  RpState saved = *c->h_state;
  RpState S = saved;
  S.phase = 0;
  S.mode = 1;
  S.ring.cur = 1;
  S.ring.next = 1;
  S.dst = 1;
  S.density_penalty = lam;
  S.wl_coef_x = wlx;
  S.wl_coef_y = wly;
  std::vector<float2> xy(ng);
  for (int k = 0; k < ng; k++) xy[k] = make_float2(cx[c->gperm[k]], cy[c->gperm[k]]);
  HIPCHK(c, hipMemcpyAsync(c->d_slp.p + (size_t)ng, xy.data(), sizeof(float2) * ng, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->d_state.p, &S, sizeof S, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(c->d_bin_inst.p, 0, sizeof(unsigned long long) * nb, s));
  HIPCHK(c, hipMemsetAsync(c->d_bin_filler.p, 0, sizeof(unsigned long long) * nb, s));
  RpDev d = c->dev;
  d.store_rho = 1;  // the density is an output here
  d.dbg_wl = c->d_dbg_wl.p;
  d.dbg_dens = c->d_dbg_dens.p;
  HIPCHK(c, rp_launch_scatter(d, s));
  if (o->bin_inst)
    HIPCHK(c, hipMemcpyAsync(o->bin_inst, c->d_bin_inst.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s));
  if (o->bin_filler)
    HIPCHK(c, hipMemcpyAsync(o->bin_filler, c->d_bin_filler.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s));
  HIPCHK(c, rp_launch_poisson(d, s, 1, 0, c->mf()));
  HIPCHK(c, rp_launch_wa_grad(d, s, 1));
  std::vector<float2> field(nb), sg(ng), wl(ng), dn(ng);
  std::vector<double> pg((size_t)RP_P_COUNT * c->h_state->num_partials_grad), pb(c->h_state->num_partials_bin);
  std::vector<long long> pn(std::max(1, c->h_state->num_partials_net));
  HIPCHK(c, hipMemcpyAsync(field.data(), c->d_field.p, sizeof(float2) * nb, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(sg.data(), c->d_sumg.p + (size_t)ng, sizeof(float2) * ng, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(wl.data(), c->d_dbg_wl.p, sizeof(float2) * ng, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(dn.data(), c->d_dbg_dens.p, sizeof(float2) * ng, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(pg.data(), c->d_part_grad.p, sizeof(double) * pg.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(pb.data(), c->d_part_bin.p, sizeof(double) * pb.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(pn.data(), c->d_part_net.p, sizeof(long long) * pn.size(), hipMemcpyDeviceToHost, s));
  if (o->density)
    HIPCHK(c, hipMemcpyAsync(o->density, c->d_rho.p, sizeof(float) * nb, hipMemcpyDeviceToHost, s));
  if (o->phi) HIPCHK(c, hipMemcpyAsync(o->phi, c->d_phi.p, sizeof(float) * nb, hipMemcpyDeviceToHost, s));
  std::vector<int4> nbox(nn);
  std::vector<float4> nsum((size_t)2 * nn);
  HIPCHK(c, hipMemcpyAsync(nbox.data(), c->d_nbox.p, sizeof(int4) * nn, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(nsum.data(), c->d_nsum.p, sizeof(float4) * 2 * nn, hipMemcpyDeviceToHost, s));
  // restore the controller state
  HIPCHK(c, hipMemcpyAsync(c->d_state.p, &saved, sizeof saved, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipStreamSynchronize(s));
  for (int b = 0; b < nb; b++) {
    if (o->ex) o->ex[b] = field[b].x;
    if (o->ey) o->ey[b] = field[b].y;
  }
  for (int k = 0; k < ng; k++) {
    int g = c->gperm[k];
    if (o->sum_grad) {
      o->sum_grad[2 * g] = sg[k].x;
      o->sum_grad[2 * g + 1] = sg[k].y;
    }
    if (o->wl_grad) {
      o->wl_grad[2 * g] = wl[k].x;
      o->wl_grad[2 * g + 1] = wl[k].y;
    }
    if (o->dens_grad) {
      o->dens_grad[2 * g] = dn[k].x;
      o->dens_grad[2 * g + 1] = dn[k].y;
    }
  }
  double sums[3] = {0, 0, 0};
  for (int k = 0; k < c->h_state->num_partials_grad; k++) {
    sums[0] += pg[(size_t)k * RP_P_COUNT + RP_P_WL];
    sums[1] += pg[(size_t)k * RP_P_COUNT + RP_P_DENS];
    sums[2] += pg[(size_t)k * RP_P_COUNT + RP_P_GRAD];
  }
  if (o->sums) memcpy(o->sums, sums, sizeof sums);
  if (o->overflow_area) {
    double ov = 0;
    for (double v : pb) ov += v;
    *o->overflow_area = (int64_t)ov;
  }
  if (o->hpwl) {
    long long h = 0;
    for (long long v : pn) h += v;
    *o->hpwl = h;
  }
  // nets back in reference order
  for (int k = 0; k < nn; k++) {
    int e = c->net_perm[k];
    if (o->net_sums) {
      float* q = o->net_sums + (size_t)8 * e;
      q[0] = nsum[2 * k].x;
      q[1] = nsum[2 * k].y;
      q[2] = nsum[2 * k].z;
      q[3] = nsum[2 * k].w;
      q[4] = nsum[2 * k + 1].x;
      q[5] = nsum[2 * k + 1].y;
      q[6] = nsum[2 * k + 1].z;
      q[7] = nsum[2 * k + 1].w;
    }
    if (o->net_box) {
      int32_t* q = o->net_box + (size_t)4 * e;
      q[0] = nbox[k].x;
      q[1] = nbox[k].y;
      q[2] = nbox[k].z;
      q[3] = nbox[k].w;
    }
  }
  return RP_OK;
}

extern "C" int rp_eval_poisson(rp_ctx* c, const float* density, float* phi, float* ex, float* ey) {
  if (!c || !c->nb_ready || !density) return RP_ERR_STATE;
  const int nb = c->nbx * c->nby;
  hipStream_t s = c->stream;
  RpState saved = *c->h_state;
  RpState S = saved;
  S.phase = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_state.p, &S, sizeof S, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->d_rho.p, density, sizeof(float) * nb, hipMemcpyHostToDevice, s));
  HIPCHK(c, rp_launch_poisson(c->dev, s, 1, 1, c->mf()));
  std::vector<float2> field(nb);
  HIPCHK(c, hipMemcpyAsync(field.data(), c->d_field.p, sizeof(float2) * nb, hipMemcpyDeviceToHost, s));
  if (phi) HIPCHK(c, hipMemcpyAsync(phi, c->d_phi.p, sizeof(float) * nb, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(c->d_state.p, &saved, sizeof saved, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipStreamSynchronize(s));
  for (int b = 0; b < nb; b++) {
    if (ex) ex[b] = field[b].x;
    if (ey) ey[b] = field[b].y;
  }
  return RP_OK;
}

// ===========================================================================
// Nesterov loop
// ===========================================================================
static int slots_per_poll(const rp_ctx* c) {
  const char* e = getenv("RP_SLOTS_PER_POLL");
  if (e && atoi(e) > 0) return atoi(e);
  int ng = c->n_g;
  return ng >= 200000 ? 8 : 32;
}

// NesterovPlace::init (src/nesterovPlace.cpp:64-175) as device launches on
// stream s: state and initial centres uploaded, the two lambda-0
// evaluations and their controller stages queued.  No host wait.
int rp_nesterov_begin(rp_ctx* c, hipStream_t s) {
  if (!c->nb_ready) {
    int r = rp_nesterov_setup(c, nullptr);
    if (r) return r;
  }
  const int ng = c->n_g;
  RpState& S = *c->h_state;
  S.phase = 0;
  S.code = 0;
  S.evals = 0;
  S.diverged = 0;
  S.mode = 1;
  S.ring.cur = 0;
  S.ring.next = 1;
  S.ring.prev = 2;
  S.ring.co_cur = 0;
  S.ring.co_next = 1;
  S.dst = S.ring.cur;
  S.density_penalty = 0.0f;  // densityPenalty_(0) during both init evaluations
  S.max_phi_coef = c->prm.max_phi_coef;
  S.step_length = 0.0f;
  S.coeff = 0.0f;
  S.iter = 0;
  c->init_xy.resize(ng);
  for (int k = 0; k < ng; k++) c->init_xy[k] = make_float2(c->g_icx[c->gperm[k]], c->g_icy[c->gperm[k]]);
  for (int slot : {0, 2})  // curSLP = prevSLP = initial density centres
    HIPCHK(c, hipMemcpyAsync(c->d_slp.p + (size_t)slot * ng, c->init_xy.data(), sizeof(float2) * ng,
                             hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->d_co.p, c->init_xy.data(), sizeof(float2) * ng, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->d_state.p, &S, sizeof S, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(c->d_bin_inst.p, 0, sizeof(unsigned long long) * c->nbx * c->nby, s));
  HIPCHK(c, hipMemsetAsync(c->d_bin_filler.p, 0, sizeof(unsigned long long) * c->nbx * c->nby, s));
  const RpDev& d = c->dev;
  // eval A at curSLP with lambda 0
  HIPCHK(c, rp_launch_scatter(d, s));
  HIPCHK(c, rp_launch_poisson(d, s, 0, 0, c->mf()));
  HIPCHK(c, rp_launch_ctrl_init(d, s, 0));
  HIPCHK(c, rp_launch_wa_grad(d, s, 0));
  HIPCHK(c, rp_launch_ctrl_init(d, s, 1));
  // eval B at prevSLP = clamp(curSLP + 100 * g)
  HIPCHK(c, rp_launch_scatter(d, s));
  HIPCHK(c, rp_launch_poisson(d, s, 0, 0, c->mf()));
  HIPCHK(c, rp_launch_wa_grad(d, s, 0));
  HIPCHK(c, rp_launch_ctrl_init(d, s, 2));
  return RP_OK;
}

// Read back the final state, records and positions after the loop stopped
// (updateDb, src/nesterovPlace.cpp:585-595: LL = dC - dD/2 of the last
// evaluated iterate).  Returns the REPL code.
int rp_nesterov_collect(rp_ctx* c, double seconds) {
  const int ng = c->n_g;
  RpState* hs = c->h_state;
  HIPCHK(c, hipMemcpy(hs, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost));
  int nrec = hs->iter;
  c->records.resize(nrec);
  if (nrec > 0) HIPCHK(c, hipMemcpy(c->records.data(), c->d_rec.p, sizeof(RpIterRec) * nrec, hipMemcpyDeviceToHost));
  std::vector<int2> ic(ng);
  HIPCHK(c, hipMemcpy(ic.data(), c->d_ic.p, sizeof(int2) * ng, hipMemcpyDeviceToHost));
  for (int g = 0; g < c->n_gi; g++) {
    int i = c->g_inst[g];
    int k = c->ginv[g];
    c->dblx[i] = ic[k].x - c->g_dhx[g];
    c->dbly[i] = ic[k].y - c->g_dhy[g];
  }
  c->placed = true;
  int code = hs->phase == 2 ? hs->code : 0;
  rp_nesterov_stats& out = c->last_stats;
  out.iterations = hs->iter;
  out.evals = hs->evals;
  out.finished = hs->phase == 1;
  out.code = code;
  out.final_overflow = hs->sum_overflow;
  out.final_step_length = hs->step_length;
  out.final_density_penalty = hs->density_penalty;
  out.diverged = hs->phase == 2 ? 1 : 0;
  out.final_hpwl = hs->prev_hpwl;
  out.seconds = seconds;
  return code;
}

namespace {

// Events and the captured graph of one run of the slot loop, released on
// every exit path.
struct LoopRes {
  hipEvent_t e0 = nullptr, e1 = nullptr, polled = nullptr;
  std::vector<hipEvent_t> kev;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  ~LoopRes() {
    for (hipEvent_t e : kev)
      if (e) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : {e0, e1, polled})
      if (e) (void)hipEventDestroy(e);
  }
};

// The doNesterovPlace loop (src/nesterovPlace.cpp:315-463) on the device:
// evaluation slots are queued in batches and the state is polled between
// batches until the device state leaves phase 0 (stop, divergence or the
// iteration cap in the state).  The host copy c->h_state holds the final
// state.
int run_slots(rp_ctx* c, hipStream_t s, LoopRes& res) {
  const RpDev& d = c->dev;
  const int K = slots_per_poll(c);
  RpState* hs = c->h_state;
  HIPCHK(c, hipEventCreateWithFlags(&res.polled, hipEventDisableTiming));
  const int max_batches = (c->prm.nesterov_place_max_iter * std::max(1, c->prm.max_back_track)) / K + 4;
  std::vector<hipEvent_t>& kev = res.kev;
  constexpr int EV = 2 * RP_NUM_EVAL_KERNELS;  // events per slot
  if (c->timing) {
    kev.assign((size_t)K * EV + 2, nullptr);  // + the sort refresh bracket
    for (auto& e : kev) HIPCHK(c, hipEventCreate(&e));
  }
  hipEvent_t* sev = c->timing ? &kev[(size_t)K * EV] : nullptr;
  unsigned launched = 0;
  HIPCHK(c, hipMemcpyAsync(hs, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  int evals_before = hs->evals;
  // tile counts left by a run that stopped after its last counting slot
  HIPCHK(c, hipMemsetAsync(c->sortb.cnt, 0, sizeof(int) * RP_SORT_TILES, s));
  // Batches after the first (sort refresh + K slots, the same launches every
  // time) are captured once into a HIP graph and replayed, unless per-kernel
  // events are recorded or the matrix-core transform (library GEMMs) is in use.
  const bool use_graph = !c->timing && !c->trace && !c->mf() && getenv("RP_NO_GRAPH") == nullptr;
  const int want_phi = c->trace ? 1 : 0;  // sumPhi of the trace needs phi
  for (int b = 0; b < max_batches && hs->phase == 0; b++) {
    if (use_graph && b > 0) {
      if (!res.gexec) {
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = rp_launch_sort_refresh(d, c->sortb, s);
        for (int k = 0; k < K && e == hipSuccess; k++)
          e = rp_launch_eval(d, s, want_phi, nullptr, nullptr, k == K - 1 ? &c->sortb : nullptr, &launched);
        const hipError_t e2 = hipStreamEndCapture(s, &res.graph);
        HIPCHK(c, e != hipSuccess ? e : e2);
        HIPCHK(c, hipGraphInstantiate(&res.gexec, res.graph, nullptr, nullptr, 0));
      }
      HIPCHK(c, hipGraphLaunch(res.gexec, s));
    } else {
      // Spatial order of the scatter from the centres counted by the last slot
      // of the previous batch (it ran: the state polled after it still runs).
      if (sev) HIPCHK(c, hipEventRecord(sev[0], s));
      if (b > 0) HIPCHK(c, rp_launch_sort_refresh(d, c->sortb, s));
      if (sev) HIPCHK(c, hipEventRecord(sev[1], s));
      for (int k = 0; k < K; k++)
        HIPCHK(c, rp_launch_eval(d, s, want_phi, c->timing ? &kev[(size_t)k * EV] : nullptr, c->mf(),
                                 k == K - 1 ? &c->sortb : nullptr, &launched));
    }
    HIPCHK(c, hipMemcpyAsync(hs, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipEventRecord(res.polled, s));
    HIPCHK(c, hipEventSynchronize(res.polled));
    if (c->timing) {
      // only slots that evaluated (the ones after the stop are no-ops)
      int ran = std::min(K, hs->evals - evals_before);
      for (int k = 0; k < ran; k++) {
        for (int q = 0; q < RP_NUM_EVAL_KERNELS; q++) {
          if (!(launched >> q & 1)) continue;
          float ms = 0;
          hipEvent_t* ev = &kev[(size_t)k * EV];
          HIPCHK(c, hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]));
          c->ktime_ms[q] += ms;
        }
      }
      c->kcount += std::max(0, ran);
      float sms = 0;
      HIPCHK(c, hipEventElapsedTime(&sms, sev[0], sev[1]));
      c->ktime_ms[RP_NUM_EVAL_KERNELS] += sms;
    }
    evals_before = hs->evals;
  }
  return 0;
}

// begin_iteration of the device controller (src/nesterovPlace.cpp:318-326)
// on the host copy of the state: the same double expressions, so the same
// floats.
void host_begin_iteration(RpState& S) {
  const float prevA = S.cur_a;
  S.prev_a = prevA;
  S.cur_a = (float)((1.0 + std::sqrt(4.0 * (double)prevA * (double)prevA + 1.0)) * 0.5);
  S.coeff = (float)(((double)prevA - 1.0) / (double)S.cur_a);
  S.bt = 0;
  S.mode = 0;
  S.dst = S.ring.next;
}

}  // namespace

extern "C" int rp_nesterov_place(rp_ctx* c, rp_nesterov_stats* st) {
  if (!c) return RP_ERR_ARG;
  hipStream_t s = c->stream;
  LoopRes res;
  HIPCHK(c, hipEventCreate(&res.e0));
  HIPCHK(c, hipEventCreate(&res.e1));
  HIPCHK(c, hipEventRecord(res.e0, s));
  int r = rp_nesterov_begin(c, s);
  if (r) return r;
  if (int q = run_slots(c, s, res)) return q;
  HIPCHK(c, hipEventRecord(res.e1, s));
  HIPCHK(c, hipEventSynchronize(res.e1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, res.e0, res.e1));
  c->np_ready = true;
  c->loop_seconds = ms * 1e-3;
  int code = rp_nesterov_collect(c, c->loop_seconds);
  if (st) *st = c->last_stats;
  return code;
}

extern "C" int rp_nesterov_init(rp_ctx* c) {
  if (!c) return RP_ERR_ARG;
  hipStream_t s = c->stream;
  LoopRes res;
  HIPCHK(c, hipEventCreate(&res.e0));
  HIPCHK(c, hipEventCreate(&res.e1));
  HIPCHK(c, hipEventRecord(res.e0, s));
  int r = rp_nesterov_begin(c, s);
  if (r) return r;
  HIPCHK(c, hipEventRecord(res.e1, s));
  HIPCHK(c, hipMemcpyAsync(c->h_state, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, res.e0, res.e1));
  c->loop_seconds = ms * 1e-3;
  c->np_ready = true;
  const RpState& S = *c->h_state;
  return S.phase == 2 ? S.code : 0;
}

extern "C" int rp_nesterov_iterate(rp_ctx* c, int32_t n_iters, rp_nesterov_stats* st) {
  if (!c || n_iters < 0) return RP_ERR_ARG;
  if (!c->np_ready) return fail(c, RP_ERR_STATE, "rp_nesterov_iterate before rp_nesterov_init");
  hipStream_t s = c->stream;
  RpState& S = *c->h_state;
  HIPCHK(c, hipMemcpy(&S, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost));
  const int cap = c->prm.nesterov_place_max_iter;
  // a pause at the bounded iteration count of an earlier call: the next
  // iteration begins now
  if (S.phase == 3 && S.iter < cap) {
    S.phase = 0;
    host_begin_iteration(S);
  }
  if (S.phase == 0 && n_iters > 0) {
    S.max_iter = std::min(cap, S.iter + n_iters);
    HIPCHK(c, hipMemcpy(c->d_state.p, &S, sizeof(RpState), hipMemcpyHostToDevice));
    LoopRes res;
    HIPCHK(c, hipEventCreate(&res.e0));
    HIPCHK(c, hipEventCreate(&res.e1));
    HIPCHK(c, hipEventRecord(res.e0, s));
    int r = run_slots(c, s, res);
    // the iteration cap of the context again (the state keeps phase 3 for a pause)
    HIPCHK(c, hipMemcpy(&S, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost));
    S.max_iter = cap;
    HIPCHK(c, hipMemcpy(c->d_state.p, &S, sizeof(RpState), hipMemcpyHostToDevice));
    if (r) return r;
    HIPCHK(c, hipEventRecord(res.e1, s));
    HIPCHK(c, hipEventSynchronize(res.e1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, res.e0, res.e1));
    c->loop_seconds += ms * 1e-3;
  }
  int code = rp_nesterov_collect(c, c->loop_seconds);
  if (st) *st = c->last_stats;
  return code;
}

extern "C" int rp_nesterov_get_state(rp_ctx* c, rp_nesterov_state* out) {
  if (!c || !out) return RP_ERR_ARG;
  if (!c->np_ready) return fail(c, RP_ERR_STATE, "rp_nesterov_get_state before rp_nesterov_init");
  const int ng = c->n_g;
  RpState& S = *c->h_state;
  HIPCHK(c, hipMemcpy(&S, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost));
  auto get = [&](float* dst, const float2* base, int slot) -> int {
    if (!dst) return 0;
    std::vector<float2> v(ng);
    HIPCHK(c, hipMemcpy(v.data(), base + (size_t)slot * ng, sizeof(float2) * ng, hipMemcpyDeviceToHost));
    for (int k = 0; k < ng; k++) {
      const int g = c->gperm[k];
      dst[2 * g] = v[k].x;
      dst[2 * g + 1] = v[k].y;
    }
    return 0;
  };
  if (int r = get(out->cur_slp, c->d_slp.p, S.ring.cur)) return r;
  if (int r = get(out->cur_slp_grad, c->d_sumg.p, S.ring.cur)) return r;
  if (int r = get(out->prev_slp, c->d_slp.p, S.ring.prev)) return r;
  if (int r = get(out->prev_slp_grad, c->d_sumg.p, S.ring.prev)) return r;
  if (int r = get(out->cur_coord, c->d_co.p, S.ring.co_cur)) return r;
  out->step_length = S.step_length;
  out->density_penalty = S.density_penalty;
  out->wl_coef_x = S.wl_coef_x;
  out->wl_coef_y = S.wl_coef_y;
  out->base_wl_coef = S.base_wl_coef;
  // a running state has begun its next iteration (cur_a advanced)
  out->cur_a = S.phase == 0 ? S.prev_a : S.cur_a;
  out->sum_overflow = S.sum_overflow;
  out->min_sum_overflow = S.min_sum_overflow;
  out->hpwl_with_min_sum_overflow = S.hpwl_with_min_sum_overflow;
  out->max_phi_coef = S.max_phi_coef;
  out->is_max_phi_coef_changed = S.is_max_phi_coef_changed;
  out->iter = S.iter;
  out->prev_hpwl = S.prev_hpwl;
  return RP_OK;
}

extern "C" int rp_nesterov_set_state(rp_ctx* c, const rp_nesterov_state* in) {
  if (!c || !in) return RP_ERR_ARG;
  if (!c->np_ready) return fail(c, RP_ERR_STATE, "rp_nesterov_set_state before rp_nesterov_init");
  if (in->iter < 0 || in->iter >= c->prm.nesterov_place_max_iter)
    return fail(c, RP_ERR_ARG, "state iteration %d outside [0, %d)", in->iter, c->prm.nesterov_place_max_iter);
  const int ng = c->n_g;
  RpState& S = *c->h_state;
  HIPCHK(c, hipMemcpy(&S, c->d_state.p, sizeof(RpState), hipMemcpyDeviceToHost));
  auto put = [&](const float* src, float2* base, int slot) -> int {
    if (!src) return 0;
    std::vector<float2> v(ng);
    for (int k = 0; k < ng; k++) {
      const int g = c->gperm[k];
      v[k] = make_float2(src[2 * g], src[2 * g + 1]);
    }
    HIPCHK(c, hipMemcpy(base + (size_t)slot * ng, v.data(), sizeof(float2) * ng, hipMemcpyHostToDevice));
    return 0;
  };
  if (int r = put(in->cur_slp, c->d_slp.p, S.ring.cur)) return r;
  if (int r = put(in->cur_slp_grad, c->d_sumg.p, S.ring.cur)) return r;
  if (int r = put(in->prev_slp, c->d_slp.p, S.ring.prev)) return r;
  if (int r = put(in->prev_slp_grad, c->d_sumg.p, S.ring.prev)) return r;
  if (int r = put(in->cur_coord, c->d_co.p, S.ring.co_cur)) return r;
  S.step_length = in->step_length;
  S.density_penalty = in->density_penalty;
  S.wl_coef_x = in->wl_coef_x;
  S.wl_coef_y = in->wl_coef_y;
  S.base_wl_coef = in->base_wl_coef;
  S.cur_a = in->cur_a;
  S.sum_overflow = in->sum_overflow;
  S.min_sum_overflow = in->min_sum_overflow;
  S.hpwl_with_min_sum_overflow = in->hpwl_with_min_sum_overflow;
  S.max_phi_coef = in->max_phi_coef;
  S.is_max_phi_coef_changed = in->is_max_phi_coef_changed;
  S.iter = in->iter;
  S.prev_hpwl = in->prev_hpwl;
  S.phase = 0;
  S.code = 0;
  S.diverged = 0;
  S.max_iter = c->prm.nesterov_place_max_iter;
  host_begin_iteration(S);
  HIPCHK(c, hipMemcpy(c->d_state.p, &S, sizeof(RpState), hipMemcpyHostToDevice));
  return RP_OK;
}

static_assert(sizeof(RpEvalRec) == sizeof(rp_eval_record), "rp_eval_record layout");

extern "C" int rp_get_eval_records(rp_ctx* c, rp_eval_record* out, int32_t cap, int32_t* count) {
  if (!c || !count) return RP_ERR_ARG;
  *count = 0;
  if (!c->trace || !c->nb_ready) return RP_OK;
  const int n = std::min<int>(c->h_state->evals, (int)c->d_evrec.n);
  *count = n;
  if (out && cap > 0)
    HIPCHK(c, hipMemcpy(out, c->d_evrec.p, sizeof(rp_eval_record) * std::min(n, cap), hipMemcpyDeviceToHost));
  return RP_OK;
}

extern "C" int rp_get_iter_records(rp_ctx* c, rp_iter_record* out, int32_t cap, int32_t* count) {
  if (!c || !count) return RP_ERR_ARG;
  *count = (int32_t)c->records.size();
  int n = std::min<int>(cap, *count);
  for (int k = 0; k < n && out; k++) {
    out[k].overflow = c->records[k].overflow;
    out[k].step_length = c->records[k].step_length;
    out[k].density_penalty = c->records[k].density_penalty;
    out[k].num_back_track = c->records[k].num_back_track;
    out[k].hpwl = c->records[k].hpwl;
  }
  return RP_OK;
}

extern "C" int rp_get_gcell_centres(rp_ctx* c, float* cx, float* cy) {
  if (!c || !cx || !cy) return RP_ERR_ARG;
  if (!c->nb_ready) return fail(c, RP_ERR_ARG, "rp_get_gcell_centres before rp_nesterov_setup");
  const int ng = c->n_g;
  std::vector<int2> ic(ng);
  HIPCHK(c, hipMemcpy(ic.data(), c->d_ic.p, sizeof(int2) * ng, hipMemcpyDeviceToHost));
  for (int g = 0; g < ng; g++) {
    int k = c->ginv[g];
    cx[g] = (float)ic[k].x;
    cy[g] = (float)ic[k].y;
  }
  return RP_OK;
}

extern "C" int rp_get_positions(rp_ctx* c, int32_t* lx, int32_t* ly) {
  if (!c || !lx || !ly) return RP_ERR_ARG;
  for (int i = 0; i < c->n_real; i++) {
    lx[i] = c->dblx[i];
    ly[i] = c->dbly[i];
  }
  return RP_OK;
}

extern "C" int rp_set_positions(rp_ctx* c, const int32_t* lx, const int32_t* ly) {
  if (!c || !lx || !ly) return RP_ERR_ARG;
  for (int i = 0; i < c->n_real; i++) {
    int w = c->iux[i] - c->ilx[i], h = c->iuy[i] - c->ily[i];
    c->ilx[i] = lx[i];
    c->ily[i] = ly[i];
    c->iux[i] = lx[i] + w;
    c->iuy[i] = ly[i] + h;
    c->dblx[i] = lx[i];
    c->dbly[i] = ly[i];
    inst_update_pins(c, i);
  }
  c->nb_ready = false;
  return RP_OK;
}

extern "C" int rp_set_kernel_timing(rp_ctx* c, int enable) {
  if (!c) return RP_ERR_ARG;
  c->timing = enable != 0;
  if (c->timing) {  // a new measurement; disabling keeps the totals for rp_get_kernel_timing
    for (double& v : c->ktime_ms) v = 0.0;
    c->kcount = 0;
  }
  return RP_OK;
}

extern "C" int rp_get_kernel_timing(rp_ctx* c, double* total_ms, int64_t* launches) {
  if (!c || !total_ms || !launches) return RP_ERR_ARG;
  for (int q = 0; q <= RP_NUM_EVAL_KERNELS; q++) total_ms[q] = c->ktime_ms[q];
  *launches = c->kcount;
  return RP_OK;
}
