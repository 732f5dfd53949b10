// rp_ctx.h -- the placement context behind rp_ctx* (private to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/replace/rp_capi.h"
#include "rp_dev_struct.h"

// Device allocations come from the device's stream-ordered memory pool
// (hipMallocAsync / hipFreeAsync on the null stream, which holds no work of
// this library), kept from one allocation to the next.  Unlike hipFree,
// releasing a buffer does not wait for the whole device, so placement
// contexts set up and torn down on host threads do not stall each other's
// queued kernels.  A buffer is released only when no queued work uses it:
// its owner synchronises the streams that use it first.
hipError_t rp_dev_malloc(void** p, size_t bytes);
void rp_dev_free(void* p);
// streams of placement contexts, reused across contexts (rp_host.cpp)
hipError_t rp_stream_acquire(int dev, hipStream_t* st);
void rp_stream_release(int dev, hipStream_t st);
// Pinned host staging blocks of the contexts (states, records, totals) come
// from a process-wide cache by size class: hipHostFree waits for the whole
// device, so a context's teardown hands its blocks back instead.
hipError_t rp_pinned_alloc(void** p, size_t bytes);
void rp_pinned_free(void* p);

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) rp_dev_free(p);
    p = nullptr;
    n = 0;
  }
};

// Matrix-core Poisson solve (rp_dct_mfma.hip): an opaque plan (GEMM
// descriptors, split-BF16 matrices, intermediates), released with the context.
struct RpMfma {
  bool ready = false;
  void* plan = nullptr;
  size_t nbytes = 0;
  RpMfmaDev view() const { return RpMfmaDev{plan}; }
  size_t bytes() const { return nbytes; }
  RpMfma() = default;
  RpMfma(const RpMfma&) = delete;
  RpMfma& operator=(const RpMfma&) = delete;
  ~RpMfma();
};

struct rp_ctx {
  rp_params prm;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  // ---- PlacerBase ----
  int core_lx = 0, core_ly = 0, core_ux = 0, core_uy = 0;
  int die_lx = 0, die_ly = 0, die_ux = 0, die_uy = 0, dbu = 0;
  int n_unplaced_ports = 0;
  int site_w = 0, site_h = 0;
  int n_real = 0;
  std::vector<int> ilx, ily, iux, iuy, mw, mh, dblx, dbly;  // incl. dummies
  std::vector<uint8_t> fixed;
  std::vector<int> place, inst2place, nonplace;
  int64_t place_area = 0, nonplace_area = 0, macro_area = 0, std_area = 0;
  int n_pin = 0, n_net = 0;
  std::vector<int> net_start, pin_inst, pin_offx, pin_offy, pin_cx, pin_cy;
  std::vector<int> bfs_order;  // rp_movable_order
  std::vector<int> pin_box;  // 4 per pin (BTerms)
  std::vector<int> ipin_start, ipins;
  // ---- initial place stats ----
  std::vector<float> ip_err;
  std::vector<int64_t> ip_hpwl;
  // ---- NesterovBase ----
  bool nb_ready = false;
  int n_g = 0, n_gi = 0, n_gf = 0;
  int filler_dx = 0, filler_dy = 0;
  int nbx = 0, nby = 0, bsx = 0, bsy = 0;
  float td = 0;
  std::vector<int> g_inst;
  std::vector<uint8_t> g_kind;
  std::vector<int> g_dx, g_dy, g_dhx, g_dhy;
  std::vector<float> g_scale, g_icx, g_icy;
  std::vector<int64_t> nonplace_bins;
  // ---- device ----
  DevBuf<int2> d_dhalf;
  DevBuf<float> d_scale, d_area, d_npin;
  DevBuf<uint8_t> d_kind;
  DevBuf<float2> d_slp, d_sumg, d_co;
  DevBuf<int2> d_ic;
  DevBuf<int> d_sord, d_srank, d_scnt, d_soff;  // spatial cell order of the scatter and its counting sort
  RpSortBufs sortb{};
  DevBuf<int> d_cp_start, d_cp_pin, d_nstart, d_pcell;
  DevBuf<int4> d_wa_rng[RP_WA_VARIANTS];
  DevBuf<int2> d_ppos;
  DevBuf<float2> d_pgrad;
  DevBuf<int4> d_nbox;
  DevBuf<float4> d_nsum;
  DevBuf<unsigned long long> d_bin_inst, d_bin_filler;
  DevBuf<float> d_bin_npf;
  DevBuf<float> d_rho, d_phi, d_t, d_y1, d_y2;
  DevBuf<float2> d_field;
  DevBuf<float2> d_twx, d_tw2x, d_twy, d_tw2y;
  DevBuf<float> d_wx, d_wx2, d_wy, d_wy2;
  DevBuf<double> d_part_grad, d_part_bin;
  DevBuf<long long> d_part_net;
  DevBuf<RpState> d_state;
  DevBuf<RpIterRec> d_rec;
  DevBuf<float2> d_dbg_wl, d_dbg_dens;
  // verbose-3 trace
  bool trace = false;
  DevBuf<RpEvalRec> d_evrec;
  DevBuf<double> d_part_gn, d_part_phi;
  DevBuf<float> d_dbg_area;
  int64_t nb_core_area = 0, nb_white = 0, nb_movable = 0, nb_total_filler = 0;
  int64_t nb_avg_area = 0, nb_ideal_bin_area = 0, nb_ideal_bin_cnt = 0;
  RpState* h_state = nullptr;  // pinned
  std::vector<float2> init_xy; // initial centres (device order), kept alive for async upload
  std::vector<int> net_perm;   // device net id -> reference net id
  std::vector<int> gperm;      // device gcell index -> reference gcell index
  std::vector<int> ginv;       // reference gcell index -> device gcell index
  RpDev dev;
  RpMfma mfma;
  RpMfmaDev mfma_view{};
  const RpMfmaDev* mf() const { return mfma.ready ? &mfma_view : nullptr; }
  // ---- per-kernel timing (optional) ----
  bool timing = false;
  double ktime_ms[RP_NUM_EVAL_KERNELS + 1] = {0};  // slot kernels, then the per-batch sort refresh
  long long kcount = 0;
  // ---- results ----
  std::vector<RpIterRec> records;
  rp_nesterov_stats last_stats;
  bool placed = false;
  bool np_ready = false;      // rp_nesterov_init ran: rp_nesterov_iterate may continue the loop
  double loop_seconds = 0.0;  // device time of the loop since rp_nesterov_init

  // device bytes held by this context (HBM admission accounting)
  size_t device_bytes() const {
    size_t b = 0;
    auto add = [&b](const auto& buf) { b += buf.n * sizeof(*buf.p); };
    add(d_dhalf); add(d_scale); add(d_area); add(d_npin); add(d_kind); add(d_slp); add(d_sumg); add(d_co);
    add(d_ic); add(d_cp_start); add(d_cp_pin); add(d_nstart); add(d_pcell); add(d_ppos); add(d_pgrad);
    add(d_nbox); add(d_nsum); add(d_bin_inst); add(d_bin_filler); add(d_bin_npf); add(d_rho); add(d_phi);
    add(d_t); add(d_y1); add(d_y2); add(d_field); add(d_twx); add(d_tw2x);
    add(d_twy); add(d_tw2y); add(d_wx); add(d_wx2); add(d_wy); add(d_wy2); add(d_part_grad); add(d_part_bin);
    add(d_part_net); add(d_state); add(d_rec); add(d_dbg_wl); add(d_dbg_dens);
    add(d_evrec); add(d_part_gn); add(d_part_phi); add(d_dbg_area);
    add(d_sord); add(d_srank); add(d_scnt); add(d_soff);
    for (const auto& w : d_wa_rng) add(w);
    return b + mfma.bytes();
  }
};

static inline int fail(rp_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIPCHK(c, x)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail((c), RP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static inline hipError_t dalloc(DevBuf<T>& b, size_t n) {
  b.release();
  if (n == 0) n = 1;
  hipError_t e = rp_dev_malloc((void**)&b.p, n * sizeof(T));
  if (e == hipSuccess) b.n = n;
  return e;
}

template <typename T>
static inline hipError_t dupload(DevBuf<T>& b, const std::vector<T>& v, hipStream_t s) {
  hipError_t e = dalloc(b, v.size());
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}


// InitialPlace B2B assembly on the device (rp_ip.hip)
struct RpIpAsm {
  int n = 0, n_net = 0, cb = 0;
  DevBuf<int> net_start, pin_inst, pin_k, pin_cx, pin_cy, icx, icy, pin_net;
  DevBuf<int4> ext;      // per net: the pins of the x/y extremes
  DevBuf<int> cnt, off;  // 4 x (n_pin + 1) per-pin counts and their exclusive offsets
  DevBuf<unsigned long long> tk[2], tk2[2];
  DevBuf<float> tv[2], tv2[2];
  DevBuf<unsigned> rk[2], rk2[2];
  DevBuf<float> rv[2], rv2[2];
  DevBuf<int> flag, seg, srow;
  DevBuf<char> temp;
  DevBuf<int> rowp[2], col[2];  // per-axis CSR and right-hand side
  DevBuf<float> val[2], b[2];
  int* h_tot = nullptr;
  std::vector<int> h_pin_k, pos;  // pin -> row (locality order), movable index -> row
  ~RpIpAsm() {
    if (h_tot) rp_pinned_free(h_tot);
  }
};
int rp_ip_asm_init(rp_ctx* c, RpIpAsm& a);
const std::vector<int>& rp_movable_order(rp_ctx* c);  // BFS locality order of movable instances
int rp_ip_assemble_a(rp_ctx* c, RpIpAsm& a, hipStream_t st);  // counts; wait for a.h_tot, then
int rp_ip_assemble_b(rp_ctx* c, RpIpAsm& a, hipStream_t st);  // the CSR and right-hand sides
// sequential host assembly of the same systems from the device's centres (rp_host.cpp)
int rp_ip_host_assemble(rp_ctx* c, RpIpAsm& a);

// The InitialPlace iterations on the device (rp_ip.hip), one context's run
// advanced phase by phase so that a driver can interleave many runs on one
// stream; rp_ip_device_loop drives a single run.  Every phase enqueues on the
// run's stream; "wait" means the caller synchronises that stream in between.
struct RpIpRun;
struct RpIpBatch;  // tables and staging of a driver that solves many runs together
int rp_ipr_begin(rp_ctx* c, hipStream_t st, bool host_asm, RpIpRun** out);  // uploads on c->stream, then waits
int rp_ipr_assemble_a(RpIpRun* r);  // wait, then
int rp_ipr_assemble_b(RpIpRun* r);  // CSR, warm start, solver inputs
// the solves of k runs (all on stream st); returns once every one has ended
int rp_ipr_solve(RpIpRun* const* runs, int k, hipStream_t st, RpIpBatch* b);
int rp_ipr_finish_a(RpIpRun* r);  // HPWL record, updateCoordi; wait, then
int rp_ipr_finish_b(RpIpRun* r, bool* stop);  // records; stop after the reference's rule
int rp_ipr_end(RpIpRun* r);                   // positions back to the host model
void rp_ipr_free(RpIpRun* r);
rp_ctx* rp_ipr_ctx(RpIpRun* r);
RpIpBatch* rp_ipb_create();
bool rp_ip_one_workgroup(int n);  // n rows: the initial place solves in one workgroup per axis
void rp_ipb_free(RpIpBatch* b);
int rp_ip_device_loop(rp_ctx* c, bool host_asm, double* t_asm, double* t_solve);
// rp_initial_place split around its iterations (rp_host.cpp)
int rp_initial_place_begin(rp_ctx* c, hipStream_t run_stream, RpIpRun** run);
int rp_initial_place_end(rp_ctx* c, RpIpRun* run);

// launch wrappers (rp_kernels.hip)
hipError_t rp_launch_eval(const RpDev& d, hipStream_t s, int want_phi, hipEvent_t* ev, const RpMfmaDev* mf,
                          const RpSortBufs* sb = nullptr, unsigned* launched = nullptr);
hipError_t rp_launch_ctrl_init(const RpDev& d, hipStream_t s, int stage);
hipError_t rp_launch_scatter(const RpDev& d, hipStream_t s);
hipError_t rp_launch_poisson(const RpDev& d, hipStream_t s, int want_phi, int from_rho, const RpMfmaDev* mf);
int rp_mfma_setup(rp_ctx* c);  // transform matrices of the matrix-core Poisson solve
hipError_t rp_launch_wa_grad(const RpDev& d, hipStream_t s, int store_nets);
hipError_t rp_launch_eval_batch(const RpBatchLaunch& b, hipStream_t s);
hipError_t rp_launch_gather_phase(const RpDev* devs, int n, int* out, hipStream_t s);

// context-level steps shared by rp_nesterov_place and the batch engine
int rp_nesterov_begin(rp_ctx* c, hipStream_t s);
int rp_nesterov_collect(rp_ctx* c, double seconds);
