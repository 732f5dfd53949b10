// rp_batch.cpp -- placement service on one GPU: continuous batching of many
// placement jobs through shared evaluation steps.
//
// No reference counterpart: the reference places one design per process
// (src/replace.cpp:73-128).  Each job here is exactly the reference's
// placement (the same per-instance kernels and controller as rp_nesterov_place);
// what is shared is the launch: every batched step runs one evaluation of
// every running job through per-kernel block tables (rp_launch_eval_batch),
// captured in a HIP graph.
//
// Scheduling loop (rp_server_run):
//   1. jobs whose arrival time has passed are admitted in arrival order while
//      their estimated device footprint fits the HBM budget and a slot is free;
//      admitted jobs are set up (PlacerBase / NesterovBase derivations and
//      upload) on host worker threads while the GPU keeps running;
//   2. set-up jobs start: their NesterovPlace::init launches go on the
//      service stream and they join the block tables;
//   3. the graph of `steps_per_poll` batched steps is replayed;
//   4. the phase word of every running job is gathered; finished jobs are
//      collected (stats, records, positions), their memory is released and
//      the budget is returned; the tables are rebuilt.
// Tile shapes: the fused step+scatter and the wirelength kernel come in
// RP_SC_VARIANTS / RP_WA_VARIANTS shapes (cells / pins per workgroup).  With
// opts.autotune the first table rebuild that sees a job size bucket times
// each shape on up to kTuneJobs of the bucket's running jobs (launches of
// that kernel alone: both are idempotent on a job's state apart from the
// bin accumulators, which are cleared again afterwards) and keeps the
// fastest for the bucket; the tables then hold one launch per shape.
// A job's arithmetic does not depend on which other jobs run beside it, so
// its result is bit-identical alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>

#include "rp_ctx.h"

namespace {

using clk = std::chrono::steady_clock;

struct Job {
  int id = 0;
  int est_cells = 0;  // movable instances (the size bucket before set-up counts fillers too)
  double arrival = 0;
  std::vector<int32_t> rows, insts, nstart, pinst, poff, pbox, ips, ipn;
  std::vector<uint8_t> fixed;
  rp_netlist nl{};
  rp_params prm{};
  rp_ctx* ctx = nullptr;
  RpIpRun* ip_run = nullptr;  // initial place handed to the batched driver
  int ip_code = 0;            // its result, when it comes back
  enum { QUEUED, SETUP, READY, RUNNING, DONE, FAILED } status = QUEUED;
  int code = 0;
  std::string err;
  double admit_s = -1, start_s = -1, done_s = -1;
  size_t reserved = 0;
  rp_nesterov_stats stats{};
  std::vector<int32_t> lx, ly;
  std::vector<rp_iter_record> rec;
};

// Device footprint estimate before setup (bytes per gcell / pin / net / bin of
// the arrays in rp_device.h); replaced by the measured size after setup.
size_t estimate_bytes(const rp_netlist& nl, const rp_params& p) {
  size_t nc = (size_t)nl.num_insts, np = (size_t)nl.num_pins, nn = (size_t)nl.num_nets;
  size_t bx = p.bin_grid_cnt_x > 0 ? p.bin_grid_cnt_x : 1024;
  size_t by = p.bin_grid_cnt_y > 0 ? p.bin_grid_cnt_y : 1024;
  return 120 * nc + 24 * np + 56 * nn + 64 * bx * by + (1 << 20);
}

// Block table with XCD-aware placement: hardware block b runs on XCD b % 8.
// Small jobs go whole onto the least-loaded XCD (their working set shares one
// L2); a job with more than 1/8 of all blocks is split into 8 contiguous
// chunks exactly as xcd_block() splits a single-instance launch.
void build_table(const std::vector<int>& counts, std::vector<int2>& tab) {
  const int X = 8;
  long total = 0;
  for (int n : counts) total += n;
  std::vector<std::vector<int2>> lists(X);
  std::vector<long> load(X, 0);
  std::vector<int> order(counts.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return counts[a] > counts[b]; });
  for (int i : order) {
    int n = counts[i];
    if (n <= 0) continue;
    if ((long)n * X > total && n >= X) {
      int per = n / X, rem = n % X;
      for (int x = 0; x < X; x++) {
        int lo = x * per + std::min(x, rem), cnt = per + (x < rem ? 1 : 0);
        for (int j = 0; j < cnt; j++) lists[x].push_back(make_int2(i, lo + j));
        load[x] += cnt;
      }
    } else {
      int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      for (int j = 0; j < n; j++) lists[x].push_back(make_int2(i, j));
      load[x] += n;
    }
  }
  size_t L = 0;
  for (auto& l : lists) L = std::max(L, l.size());
  tab.assign(L * X, make_int2(-1, 0));
  for (int x = 0; x < X; x++)
    for (size_t j = 0; j < lists[x].size(); j++) tab[j * X + x] = lists[x][j];
}

constexpr int kTuneJobs = 16;   // jobs per bucket the tuner times
constexpr int kTuneMinJobs = 4;  // running jobs a bucket needs before it is tuned
constexpr int kTuneReps = 5;     // timed launches per shape (the fastest counts: other
                                 // streams' set-up kernels can share the GPU meanwhile)
constexpr int kIpDrivers = 2;     // batched initial place: driver threads, each with its own rounds
constexpr int kIpAux = 4;        // batched initial place: streams (and host threads) for the
                                 // per-run assembly and records (the hardware queues per process)

int size_bucket(int ncells) {
  if (ncells < 2000) return 0;
  if (ncells < 8000) return 1;
  if (ncells < 32000) return 2;
  if (ncells < 128000) return 3;
  return 4;
}

}  // namespace

// One driver of the batched initial place: its queue of runs, host thread,
// solve stream, aux streams and the tables of its shared solves.
struct IpDriver {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job*> queue;
  std::thread th;
  hipStream_t stream = nullptr;
  hipStream_t aux[kIpAux] = {};
  hipEvent_t asm_done[kIpAux] = {};
  hipEvent_t solved = nullptr;
  RpIpBatch* ipb = nullptr;
};

struct rp_server {
  rp_server_opts opt{};
  rp_tile_choice tiles[RP_NUM_SIZE_BUCKETS]{};
  DevBuf<int2> d_tune_tab;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<Job>> jobs;
  // tables of the running set
  DevBuf<RpDev> d_devs;
  DevBuf<int2> d_tab[RP_NUM_EVAL_KERNELS];
  DevBuf<int2> d_tab_sc[RP_SC_VARIANTS], d_tab_wa[RP_WA_VARIANTS];
  DevBuf<int> d_phase;
  int* h_phase = nullptr;
  size_t h_phase_cap = 0;
  RpBatchLaunch launch{};
  std::vector<int> running;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // budget
  size_t reserved = 0;
  // setup workers
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> setup_queue;
  bool stop = false;
  std::vector<std::thread> workers;
  // batched initial place: set-up workers hand their runs to the drivers
  // (job % kIpDrivers); a driver's run has its assembly and record work on
  // aux[job / kIpDrivers % kIpAux] (one host thread per stream issues them)
  // and the driver's shared solves on its stream.  Two drivers keep the GPU
  // busy with one's assemblies while the other solves.
  IpDriver ipd[kIpDrivers];
  bool batched_ip = false;
  clk::time_point t0{};  // start of rp_server_run (log times)
  rp_server_stats stats{};
};


static int sfail(rp_server* s, int code, const char* msg) {
  s->err = msg;
  return code;
}

#define SCHK(s, x)                                                         \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      (s)->err = std::string(#x) + ": " + hipGetErrorString(e_);           \
      return RP_ERR_HIP;                                                   \
    }                                                                      \
  } while (0)

extern "C" void rp_server_opts_default(rp_server_opts* o) {
  o->hbm_budget_bytes = (int64_t)200 << 30;
  o->max_active = 512;
  o->setup_threads = 16;
  o->steps_per_poll = 8;
  o->use_graph = 1;
  o->autotune = 1;
  o->batched_ip = 2;
}

static void tiles_default(rp_server* s) {
  for (int b = 0; b < RP_NUM_SIZE_BUCKETS; b++) {
    rp_tile_choice& t = s->tiles[b];
    memset(&t, 0, sizeof t);
    t.bucket = b;
    t.scatter_cells = 1024;
    t.wa_pins = 256 << RP_WA_DEFAULT_VARIANT;
  }
}

static int variant_of_cells(int cells) { return cells <= 256 ? 0 : cells <= 512 ? 1 : 2; }
static int variant_of_pins(int pins) { return pins <= 256 ? 0 : pins <= 512 ? 1 : 2; }

// Times every tile shape of the scatter and wirelength kernels on up to
// kTuneJobs running jobs of bucket b (their device views in devs) and keeps
// the fastest.  Runs on the service stream between batched steps.
static int tune_bucket(rp_server* s, int b, const std::vector<int>& idx, const std::vector<RpDev>& devs) {
  rp_tile_choice& t = s->tiles[b];
  const int n = std::min<int>(kTuneJobs, (int)idx.size());
  std::vector<RpDev> sample(n);
  for (int i = 0; i < n; i++) sample[i] = devs[idx[i]];
  DevBuf<RpDev> d_sample;
  SCHK(s, dupload(d_sample, sample, s->stream));
  hipEvent_t e0, e1;
  SCHK(s, hipEventCreate(&e0));
  SCHK(s, hipEventCreate(&e1));
  auto time_variant = [&](int kernel, int v, float& us) -> int {
    std::vector<int> cnt(n);
    for (int i = 0; i < n; i++)
      cnt[i] = kernel == RP_K_SCATTER ? (sample[i].nc + (256 << v) - 1) / (256 << v) : sample[i].nwa_v[v];
    std::vector<int2> tab;
    build_table(cnt, tab);
    if (tab.empty()) {
      us = 0.0f;
      return RP_OK;
    }
    std::vector<RpDev> dv = sample;
    for (auto& d : dv) {  // the wirelength shape needs its own packing
      d.nwa = d.nwa_v[kernel == RP_K_WA ? v : RP_WA_DEFAULT_VARIANT];
      d.wa_rng = d.wa_rng_v[kernel == RP_K_WA ? v : RP_WA_DEFAULT_VARIANT];
    }
    SCHK(s, hipMemcpyAsync(d_sample.p, dv.data(), sizeof(RpDev) * n, hipMemcpyHostToDevice, s->stream));
    if (s->d_tune_tab.n < tab.size()) SCHK(s, dalloc(s->d_tune_tab, 2 * tab.size()));
    SCHK(s, hipMemcpyAsync(s->d_tune_tab.p, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice, s->stream));
    SCHK(s, rp_launch_batch_variant(kernel, v, d_sample.p, s->d_tune_tab.p, (int)tab.size(), s->stream));  // warm
    us = 1e30f;
    for (int r = 0; r < kTuneReps; r++) {
      SCHK(s, hipEventRecord(e0, s->stream));
      SCHK(s, rp_launch_batch_variant(kernel, v, d_sample.p, s->d_tune_tab.p, (int)tab.size(), s->stream));
      SCHK(s, hipEventRecord(e1, s->stream));
      SCHK(s, hipEventSynchronize(e1));
      float ms = 0;
      SCHK(s, hipEventElapsedTime(&ms, e0, e1));
      us = std::min(us, ms * 1e3f);
    }
    return RP_OK;
  };
  int best_sc = 0, best_wa = 0;
  for (int v = 0; v < RP_SC_VARIANTS; v++) {
    if (int r = time_variant(RP_K_SCATTER, v, t.scatter_us[v])) return r;
    if (t.scatter_us[v] < t.scatter_us[best_sc]) best_sc = v;
  }
  for (int v = 0; v < RP_WA_VARIANTS; v++) {
    if (int r = time_variant(RP_K_WA, v, t.wa_us[v])) return r;
    if (t.wa_us[v] < t.wa_us[best_wa]) best_wa = v;
  }
  // the timed scatters added into the bin accumulators: clear them again
  for (int i = 0; i < n; i++) {
    const RpDev& d = sample[i];
    size_t nb = (size_t)d.nbx * d.nby;
    SCHK(s, hipMemsetAsync(d.bin_inst, 0, sizeof(unsigned long long) * nb, s->stream));
    SCHK(s, hipMemsetAsync(d.bin_filler, 0, sizeof(unsigned long long) * nb, s->stream));
  }
  SCHK(s, hipStreamSynchronize(s->stream));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  t.tuned = 1;
  t.jobs_sampled = n;
  t.scatter_cells = 256 << best_sc;
  t.wa_pins = 256 << best_wa;
  return RP_OK;
}

static void setup_worker(rp_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    int id;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->setup_queue.empty(); });
      if (s->stop && s->setup_queue.empty()) return;
      id = s->setup_queue.front();
      s->setup_queue.pop_front();
    }
    Job& j = *s->jobs[id];
    int r = 0;
    rp_ctx* c = nullptr;
    if (j.ip_run) {  // back from the batched initial place
      c = rp_ipr_ctx(j.ip_run);
      r = j.ip_code;
      rp_ipr_free(j.ip_run);
      j.ip_run = nullptr;
    } else {
      r = rp_create(&j.nl, &j.prm, &c);
      // batched_ip 2: only the systems of the grid-wide solver go to the rounds
      if (r == 0 && s->batched_ip && (s->opt.batched_ip != 2 || !rp_ip_one_workgroup((int)c->place.size()))) {
        IpDriver& d = s->ipd[j.id % kIpDrivers];
        RpIpRun* run = nullptr;
        r = rp_initial_place_begin(c, d.aux[j.id / kIpDrivers % kIpAux], &run);
        if (r == 0 && run) {  // the driver iterates it and queues the job here again
          j.ip_run = run;
          std::lock_guard<std::mutex> lk(d.mu);
          d.queue.push_back(&j);
          d.cv.notify_one();
          continue;
        }
        if (r == 0) r = rp_initial_place_end(c, nullptr);
      } else if (r == 0) {
        r = rp_initial_place(c, nullptr);
      }
    }
    if (r == 0) r = rp_nesterov_setup(c, nullptr);
    std::lock_guard<std::mutex> lk(s->mu);
    j.ctx = c;
    if (r != 0) {
      j.code = r;
      j.err = c ? rp_last_error(c) : "rp_create failed";
      j.status = Job::FAILED;
    } else {
      j.status = Job::READY;
    }
  }
}

// The batched initial place: every round runs one IP iteration of every job
// handed in (assembly halves back to back, the solves in shared launches, the
// records after one wait), so a round costs about what one job's iteration
// costs while all jobs advance.  A finished job, its positions written to the
// host model, goes back to the set-up queue for the rest of its set-up.
static void ip_worker(rp_server* s, IpDriver* dp) {
  IpDriver& d = *dp;
  (void)hipSetDevice(s->device);
  std::vector<Job*> active;
  const bool timing = getenv("RP_IP_TIMING") != nullptr;
  auto ms = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
  for (int round = 0;; round++) {
    {
      std::unique_lock<std::mutex> lk(d.mu);
      if (active.empty()) d.cv.wait(lk, [&] { return s->stop || !d.queue.empty(); });
      if (active.empty() && d.queue.empty()) return;  // stopping
      while (!d.queue.empty()) {
        active.push_back(d.queue.front());
        d.queue.pop_front();
      }
    }
    hipStream_t st = d.stream;
    const auto t0 = clk::now();
    // fn(job) for every active job: one host thread per aux stream issues the
    // launches of that stream's runs
    auto per_stream = [&](auto&& fn) {
      std::thread th[kIpAux];
      for (int h = 0; h < kIpAux; h++)
        th[h] = std::thread([&, h] {
          (void)hipSetDevice(s->device);
          for (Job* j : active)
            if (j->id / kIpDrivers % kIpAux == h) fn(j);
        });
      for (auto& t : th) t.join();
    };
    auto sync_aux = [&] {
      bool ok = true;
      for (int h = 0; h < kIpAux; h++) ok = hipStreamSynchronize(d.aux[h]) == hipSuccess && ok;
      return ok;
    };
    per_stream([](Job* j) {
      if (!j->ip_code) j->ip_code = rp_ipr_assemble_a(j->ip_run);
    });
    const double t_la = ms(t0);
    const bool synced = sync_aux();
    const double t_a = ms(t0);
    per_stream([&](Job* j) {
      if (!synced && !j->ip_code) j->ip_code = RP_ERR_HIP;
      if (!j->ip_code) j->ip_code = rp_ipr_assemble_b(j->ip_run);
    });
    std::vector<RpIpRun*> live;
    std::vector<Job*> lj;
    for (Job* j : active)
      if (!j->ip_code) {
        live.push_back(j->ip_run);
        lj.push_back(j);
      }
    // the shared solves wait for every run's assembly; the records for the solves
    bool evok = true;
    for (int h = 0; h < kIpAux; h++) {
      evok = hipEventRecord(d.asm_done[h], d.aux[h]) == hipSuccess && evok;
      evok = hipStreamWaitEvent(st, d.asm_done[h], 0) == hipSuccess && evok;
    }
    const double t_b = ms(t0);
    if (int r = evok ? rp_ipr_solve(live.data(), (int)live.size(), st, d.ipb) : RP_ERR_HIP)
      for (Job* j : lj) j->ip_code = r;
    evok = hipEventRecord(d.solved, st) == hipSuccess;
    for (int h = 0; h < kIpAux; h++) evok = hipStreamWaitEvent(d.aux[h], d.solved, 0) == hipSuccess && evok;
    const double t_s = ms(t0);
    per_stream([&](Job* j) {
      if (!evok && !j->ip_code) j->ip_code = RP_ERR_HIP;
      if (!j->ip_code) j->ip_code = rp_ipr_finish_a(j->ip_run);
    });
    const bool synced2 = sync_aux();
    if (timing)
      fprintf(stderr, "[rp] ip driver %d round %d at %.3f s: %zu runs, assembly launches %.1f ms, counts %.1f ms, "
              "CSR %.1f ms, solves %.1f ms, records %.1f ms\n", (int)(dp - s->ipd), round,
              std::chrono::duration<double>(t0 - s->t0).count(), live.size(), t_la, t_a - t_la, t_b - t_a,
              t_s - t_b, ms(t0) - t_s);
    std::vector<Job*> keep, back;
    for (Job* j : active) {
      bool stop = false;
      if (!synced2 && !j->ip_code) j->ip_code = RP_ERR_HIP;
      if (!j->ip_code) j->ip_code = rp_ipr_finish_b(j->ip_run, &stop);
      if (!j->ip_code && stop) j->ip_code = rp_initial_place_end(rp_ipr_ctx(j->ip_run), j->ip_run);
      if (j->ip_code || stop)
        back.push_back(j);
      else
        keep.push_back(j);
    }
    if (!back.empty()) {
      std::lock_guard<std::mutex> lk(s->mu);
      for (Job* j : back) s->setup_queue.push_back(j->id);
    }
    for (size_t k = 0; k < back.size(); k++) s->cv.notify_one();
    active.swap(keep);
  }
}

extern "C" int rp_server_create(const rp_server_opts* opt, rp_server** out) {
  if (!out) return RP_ERR_ARG;
  std::unique_ptr<rp_server> s(new rp_server());
  if (opt)
    s->opt = *opt;
  else
    rp_server_opts_default(&s->opt);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    *out = s.release();
    return sfail(*out, RP_ERR_NO_DEVICE, "no HIP device visible");
  }
  (void)hipGetDevice(&s->device);
  // the library's code objects are loaded here, on one thread, before the
  // set-up threads make their first launches concurrently
  if (rp_kernels_load_code() != hipSuccess || rp_ip_load_code() != hipSuccess) {
    *out = s.release();
    return sfail(*out, RP_ERR_HIP, "loading the kernels' code objects failed");
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    *out = s.release();
    return sfail(*out, RP_ERR_HIP, "hipStreamCreate failed");
  }
  tiles_default(s.get());
  rp_server* raw = s.get();
  if (s->opt.batched_ip) {
    bool ok = true;
    for (IpDriver& d : s->ipd) {
      ok = ok && hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&d.solved, hipEventDisableTiming) == hipSuccess;
      for (int h = 0; h < kIpAux && ok; h++)
        ok = hipStreamCreateWithFlags(&d.aux[h], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&d.asm_done[h], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
      *out = s.release();
      return sfail(*out, RP_ERR_HIP, "hipStreamCreate / hipEventCreate failed");
    }
    s->batched_ip = true;
    for (IpDriver& d : s->ipd) {
      d.ipb = rp_ipb_create();
      d.th = std::thread(ip_worker, raw, &d);
    }
  }
  int nt = std::max(1, s->opt.setup_threads);
  for (int k = 0; k < nt; k++) s->workers.emplace_back(setup_worker, raw);
  *out = s.release();
  return RP_OK;
}

extern "C" int rp_server_submit(rp_server* s, const rp_netlist* nl, const rp_params* p, double arrival_s,
                                int32_t* job) {
  if (!s || !nl || !p) return RP_ERR_ARG;
  if (p->poisson_transform != 0) return sfail(s, RP_ERR_ARG, "the placement service runs the FFT Poisson transform only");
  std::unique_ptr<Job> j(new Job());
  j->id = (int)s->jobs.size();
  j->arrival = arrival_s;
  j->prm = *p;
  auto cp = [](std::vector<int32_t>& v, const int32_t* a, size_t n) {
    if (a) v.assign(a, a + n);
  };
  cp(j->rows, nl->row_box, 4 * (size_t)nl->num_rows);
  cp(j->insts, nl->inst_box, 4 * (size_t)nl->num_insts);
  j->fixed.assign(nl->inst_fixed, nl->inst_fixed + nl->num_insts);
  cp(j->nstart, nl->net_pin_start, (size_t)nl->num_nets + 1);
  cp(j->pinst, nl->pin_inst, (size_t)nl->num_pins);
  cp(j->poff, nl->pin_off, 2 * (size_t)nl->num_pins);
  cp(j->pbox, nl->pin_box, 4 * (size_t)nl->num_pins);
  if (nl->inst_pin_start && nl->inst_pins) {
    cp(j->ips, nl->inst_pin_start, (size_t)nl->num_insts + 1);
    cp(j->ipn, nl->inst_pins, (size_t)nl->inst_pin_start[nl->num_insts]);
  }
  j->nl = *nl;
  j->nl.row_box = j->rows.data();
  j->nl.inst_box = j->insts.data();
  j->nl.inst_fixed = j->fixed.data();
  j->nl.net_pin_start = j->nstart.data();
  j->nl.pin_inst = j->pinst.data();
  j->nl.pin_off = j->poff.data();
  j->nl.pin_box = nl->pin_box ? j->pbox.data() : nullptr;
  j->nl.inst_pin_start = j->ips.empty() ? nullptr : j->ips.data();
  j->nl.inst_pins = j->ips.empty() ? nullptr : j->ipn.data();
  for (uint8_t f : j->fixed) j->est_cells += f ? 0 : 1;
  if (job) *job = j->id;
  s->jobs.push_back(std::move(j));
  return RP_OK;
}

static int rebuild(rp_server* s) {
  const int n = (int)s->running.size();
  const bool timing = getenv("RP_IP_TIMING") != nullptr;
  auto tr0 = clk::now();
  auto lap = [&](const char* what) {
    const double t = std::chrono::duration<double, std::milli>(clk::now() - tr0).count();
    if (timing && t > 10.0)
      fprintf(stderr, "[rp] slow rebuild step at %.3f s: %s %.1f ms (%d jobs)\n",
              std::chrono::duration<double>(clk::now() - s->t0).count(), what, t, n);
    tr0 = clk::now();
  };
  std::vector<RpDev> devs(n);
  std::vector<int> cnt[RP_NUM_EVAL_KERNELS];
  std::vector<int> cnt_sc[RP_SC_VARIANTS], cnt_wa[RP_WA_VARIANTS];
  size_t lb = 0, lc = 0, lr = 0;
  SCHK(s, hipStreamSynchronize(s->stream));
  lap("stream sync");
  if (s->opt.autotune) {
    std::vector<int> in_bucket[RP_NUM_SIZE_BUCKETS];
    for (int i = 0; i < n; i++) {
      devs[i] = s->jobs[s->running[i]]->ctx->dev;
      in_bucket[size_bucket(devs[i].nc)].push_back(i);
    }
    // a bucket is tuned once it has kTuneMinJobs running jobs, or when no
    // job of its size is still queued or being set up
    int waiting[RP_NUM_SIZE_BUCKETS] = {0};
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (auto& jp : s->jobs)
        if (jp->status == Job::QUEUED || jp->status == Job::SETUP || jp->status == Job::READY)
          waiting[size_bucket(jp->est_cells)]++;
    }
    for (int b = 0; b < RP_NUM_SIZE_BUCKETS; b++)
      if (!s->tiles[b].tuned && !in_bucket[b].empty() &&
          ((int)in_bucket[b].size() >= kTuneMinJobs || waiting[b] == 0))
        if (int r = tune_bucket(s, b, in_bucket[b], devs)) return r;
    lap("tuner");
  }
  for (int i = 0; i < n; i++) {
    RpDev d = s->jobs[s->running[i]]->ctx->dev;
    const rp_tile_choice& t = s->tiles[size_bucket(d.nc)];
    const int vs = variant_of_cells(t.scatter_cells), vw = variant_of_pins(t.wa_pins);
    d.nwa = d.nwa_v[vw];
    d.wa_rng = d.wa_rng_v[vw];
    devs[i] = d;
    cnt[RP_K_STEP].push_back(0);     // the batched scatter kernel steps its own cells
    cnt[RP_K_SCATTER].push_back(0);  // launched per tile shape below
    for (int v = 0; v < RP_SC_VARIANTS; v++) cnt_sc[v].push_back(v == vs ? (d.nc + (256 << v) - 1) / (256 << v) : 0);
    for (int v = 0; v < RP_WA_VARIANTS; v++) cnt_wa[v].push_back(v == vw ? d.nwa : 0);
    cnt[RP_K_PIN_POS].push_back(0);  // pin centres: computed by the wirelength kernel
    cnt[RP_K_BINFIN_ROWDCT].push_back(d.nby / 2);
    cnt[RP_K_MFMA_POISSON].push_back(0);  // batched instances use the FFT solve
    cnt[RP_K_COL_SPECTRAL].push_back(d.nbx / 2);
    cnt[RP_K_ROW_INVERSE].push_back(d.nby / 2);
    cnt[RP_K_WA].push_back(0);  // launched per tile shape
    cnt[RP_K_GRAD].push_back((d.nc + RP_BLOCK - 1) / RP_BLOCK);
    cnt[RP_K_CONTROL].push_back(1);
    lb = std::max(lb, rp_lds_binfin(d.nbx));
    lc = std::max(lc, rp_lds_col(d.nby));
    lr = std::max(lr, rp_lds_rinv(d.nbx));
  }
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
  SCHK(s, hipStreamSynchronize(s->stream));
  if (s->d_devs.n < (size_t)std::max(1, n)) SCHK(s, dalloc(s->d_devs, std::max(1, 2 * n)));
  if (n) SCHK(s, hipMemcpyAsync(s->d_devs.p, devs.data(), sizeof(RpDev) * n, hipMemcpyHostToDevice, s->stream));
  RpBatchLaunch& L = s->launch;
  L.devs = s->d_devs.p;
  std::vector<int2> tabs[RP_NUM_EVAL_KERNELS];  // host sources live until the sync below
  for (int k = 0; k < RP_NUM_EVAL_KERNELS; k++) {
    std::vector<int2>& tab = tabs[k];
    if (k == RP_K_CONTROL) {
      for (int i = 0; i < n; i++) tab.push_back(make_int2(i, 0));
    } else {
      build_table(cnt[k], tab);
    }
    if (s->d_tab[k].n < std::max<size_t>(1, tab.size())) SCHK(s, dalloc(s->d_tab[k], std::max<size_t>(1, 2 * tab.size())));
    if (!tab.empty())
      SCHK(s, hipMemcpyAsync(s->d_tab[k].p, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice, s->stream));
    L.tab[k] = s->d_tab[k].p;
    L.nblk[k] = (int)tab.size();
  }
  std::vector<int2> tabs_sc[RP_SC_VARIANTS], tabs_wa[RP_WA_VARIANTS];
  auto upload_variant = [&](const std::vector<int>& c, std::vector<int2>& tab, DevBuf<int2>& buf, const int2*& p,
                            int& nb) -> int {
    build_table(c, tab);
    if (buf.n < std::max<size_t>(1, tab.size())) SCHK(s, dalloc(buf, std::max<size_t>(1, 2 * tab.size())));
    if (!tab.empty())
      SCHK(s, hipMemcpyAsync(buf.p, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice, s->stream));
    p = buf.p;
    nb = (int)tab.size();
    return RP_OK;
  };
  for (int v = 0; v < RP_SC_VARIANTS; v++)
    if (int r = upload_variant(cnt_sc[v], tabs_sc[v], s->d_tab_sc[v], L.tab_sc[v], L.nblk_sc[v])) return r;
  for (int v = 0; v < RP_WA_VARIANTS; v++)
    if (int r = upload_variant(cnt_wa[v], tabs_wa[v], s->d_tab_wa[v], L.tab_wa[v], L.nblk_wa[v])) return r;
  L.lds_binfin = lb;
  L.lds_col = lc;
  L.lds_rinv = lr;
  if (s->d_phase.n < (size_t)std::max(1, n)) SCHK(s, dalloc(s->d_phase, std::max(1, 2 * n)));
  if (s->h_phase_cap < (size_t)std::max(1, n)) {
    if (s->h_phase) (void)hipHostFree(s->h_phase);
    s->h_phase_cap = (size_t)std::max(1, 2 * n);
    SCHK(s, hipHostMalloc((void**)&s->h_phase, sizeof(int) * s->h_phase_cap, hipHostMallocDefault));
  }
  SCHK(s, hipStreamSynchronize(s->stream));
  lap("tables and uploads");
  if (s->opt.use_graph && n > 0) {
    SCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
    for (int k = 0; k < std::max(1, s->opt.steps_per_poll); k++) {
      hipError_t e = rp_launch_eval_batch(s->launch, s->stream);
      if (e != hipSuccess) {
        hipGraph_t g;
        (void)hipStreamEndCapture(s->stream, &g);
        s->err = hipGetErrorString(e);
        return RP_ERR_HIP;
      }
    }
    SCHK(s, hipStreamEndCapture(s->stream, &s->graph));
    SCHK(s, hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
    lap("graph capture");
  }
  s->stats.rebuilds++;
  return RP_OK;
}

extern "C" int rp_server_run(rp_server* s) {
  if (!s) return RP_ERR_ARG;
  const auto t0 = clk::now();
  s->t0 = t0;
  auto now_s = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
  std::vector<int> order(s->jobs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return s->jobs[a]->arrival < s->jobs[b]->arrival; });
  const size_t budget = (size_t)std::max<int64_t>(0, s->opt.hbm_budget_bytes);
  size_t next = 0, done = 0, in_flight = 0;
  size_t total = s->jobs.size();
  bool dirty = false;
  while (done < total) {
    double t = now_s();
    // 1. admission in arrival order (FIFO, no overtaking)
    {
      std::lock_guard<std::mutex> lk(s->mu);
      while (next < order.size()) {
        Job& j = *s->jobs[order[next]];
        if (j.arrival > t) break;
        size_t est = estimate_bytes(j.nl, j.prm);
        if (in_flight > 0 && (s->reserved + est > budget || (int)in_flight >= s->opt.max_active)) break;
        j.reserved = est;
        s->reserved += est;
        j.status = Job::SETUP;
        j.admit_s = t;
        in_flight++;
        s->setup_queue.push_back(j.id);
        next++;
      }
    }
    s->cv.notify_all();
    // 2. start jobs whose setup finished
    {
      std::vector<int> ready, failed;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto& jp : s->jobs) {
          if (jp->status == Job::READY) ready.push_back(jp->id);
          if (jp->status == Job::FAILED && jp->done_s < 0) failed.push_back(jp->id);
        }
      }
      for (int id : failed) {
        Job& j = *s->jobs[id];
        j.done_s = now_s();
        if (j.ctx) rp_destroy(j.ctx);
        j.ctx = nullptr;
        s->reserved -= j.reserved;
        in_flight--;
        done++;
      }
      for (int id : ready) {
        Job& j = *s->jobs[id];
        size_t actual = j.ctx->device_bytes();
        s->reserved = s->reserved - j.reserved + actual;
        j.reserved = actual;
        int r = rp_nesterov_begin(j.ctx, s->stream);
        if (r) {
          j.code = r;
          j.err = rp_last_error(j.ctx);
          j.status = Job::FAILED;
          continue;
        }
        j.status = Job::RUNNING;
        j.start_s = now_s();
        s->running.push_back(id);
        dirty = true;
      }
    }
    if (s->running.empty()) {
      std::this_thread::sleep_for(std::chrono::microseconds(100));
      continue;
    }
    if (dirty) {
      int r = rebuild(s);
      if (r) return r;
      dirty = false;
    }
    // 3. batched steps
    const int K = std::max(1, s->opt.steps_per_poll);
    const auto tst = clk::now();
    if (s->gexec) {
      SCHK(s, hipGraphLaunch(s->gexec, s->stream));
    } else {
      for (int k = 0; k < K; k++) SCHK(s, rp_launch_eval_batch(s->launch, s->stream));
    }
    s->stats.steps += K;
    // 4. phases; retire finished jobs
    const int n = (int)s->running.size();
    SCHK(s, rp_launch_gather_phase(s->d_devs.p, n, s->d_phase.p, s->stream));
    SCHK(s, hipMemcpyAsync(s->h_phase, s->d_phase.p, sizeof(int) * n, hipMemcpyDeviceToHost, s->stream));
    SCHK(s, hipStreamSynchronize(s->stream));
    if (getenv("RP_IP_TIMING")) {
      const double t = std::chrono::duration<double, std::milli>(clk::now() - tst).count();
      if (t > 30.0) fprintf(stderr, "[rp] slow batched steps at %.3f s: %.1f ms (%d jobs)\n", now_s(), t, n);
    }
    std::vector<int> keep;
    for (int i = 0; i < n; i++) {
      Job& j = *s->jobs[s->running[i]];
      if (s->h_phase[i] == 0) {
        keep.push_back(j.id);
        continue;
      }
      double td = now_s();
      j.code = rp_nesterov_collect(j.ctx, td - j.start_s);
      j.stats = j.ctx->last_stats;
      j.rec.resize(j.ctx->records.size());
      for (size_t k = 0; k < j.rec.size(); k++) {
        const RpIterRec& r = j.ctx->records[k];
        j.rec[k] = rp_iter_record{r.overflow, r.step_length, r.density_penalty, r.num_back_track, r.hpwl};
      }
      j.lx.resize(j.nl.num_insts);
      j.ly.resize(j.nl.num_insts);
      rp_get_positions(j.ctx, j.lx.data(), j.ly.data());
      const auto tdd = clk::now();
      rp_destroy(j.ctx);
      if (getenv("RP_IP_TIMING")) {
        const double t = std::chrono::duration<double, std::milli>(clk::now() - tdd).count();
        if (t > 10.0) fprintf(stderr, "[rp] slow context release at %.3f s: %.1f ms\n", now_s(), t);
      }
      j.ctx = nullptr;
      s->reserved -= j.reserved;
      j.done_s = td;
      j.status = Job::DONE;
      in_flight--;
      done++;
      dirty = true;
    }
    s->running.swap(keep);
  }
  s->stats.wall_s = now_s();
  return RP_OK;
}

extern "C" int rp_server_job(rp_server* s, int32_t id, rp_job_result* out) {
  if (!s || !out || id < 0 || id >= (int)s->jobs.size()) return RP_ERR_ARG;
  const Job& j = *s->jobs[id];
  memset(out, 0, sizeof *out);
  out->stats = j.stats;
  out->code = j.code;
  out->done = j.status == Job::DONE;
  out->arrival_s = j.arrival;
  out->admit_s = j.admit_s;
  out->start_s = j.start_s;
  out->done_s = j.done_s;
  out->device_bytes = (int64_t)j.reserved;
  return RP_OK;
}

extern "C" int rp_server_job_positions(rp_server* s, int32_t id, int32_t* lx, int32_t* ly) {
  if (!s || id < 0 || id >= (int)s->jobs.size() || !lx || !ly) return RP_ERR_ARG;
  const Job& j = *s->jobs[id];
  if (j.status != Job::DONE) return RP_ERR_STATE;
  std::copy(j.lx.begin(), j.lx.end(), lx);
  std::copy(j.ly.begin(), j.ly.end(), ly);
  return RP_OK;
}

extern "C" int rp_server_job_records(rp_server* s, int32_t id, rp_iter_record* out, int32_t cap, int32_t* count) {
  if (!s || id < 0 || id >= (int)s->jobs.size() || !count) return RP_ERR_ARG;
  const Job& j = *s->jobs[id];
  *count = (int32_t)j.rec.size();
  for (int k = 0; k < std::min<int>(cap, *count) && out; k++) out[k] = j.rec[k];
  return RP_OK;
}

extern "C" int rp_server_get_stats(rp_server* s, rp_server_stats* out) {
  if (!s || !out) return RP_ERR_ARG;
  *out = s->stats;
  return RP_OK;
}

extern "C" int rp_server_set_tiles(rp_server* s, int32_t bucket, int32_t scatter_cells, int32_t wa_pins) {
  if (!s || bucket < 0 || bucket >= RP_NUM_SIZE_BUCKETS) return RP_ERR_ARG;
  auto ok = [](int v) { return v == 256 || v == 512 || v == 1024; };
  if (!ok(scatter_cells) || !ok(wa_pins)) return sfail(s, RP_ERR_ARG, "tile sizes are 256, 512 or 1024");
  rp_tile_choice& t = s->tiles[bucket];
  t.scatter_cells = scatter_cells;
  t.wa_pins = wa_pins;
  t.tuned = 1;  // pinned: the tuner leaves this bucket alone
  t.jobs_sampled = 0;
  return RP_OK;
}

extern "C" int rp_server_tuning(rp_server* s, rp_tile_choice* out, int32_t cap, int32_t* count) {
  if (!s || !count) return RP_ERR_ARG;
  *count = RP_NUM_SIZE_BUCKETS;
  for (int b = 0; b < std::min<int>(cap, RP_NUM_SIZE_BUCKETS) && out; b++) out[b] = s->tiles[b];
  return RP_OK;
}

extern "C" const char* rp_server_last_error(rp_server* s) { return s ? s->err.c_str() : "null server"; }

extern "C" void rp_server_destroy(rp_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto& t : s->workers) t.join();
  for (IpDriver& d : s->ipd) {
    if (d.th.joinable()) {
      {
        std::lock_guard<std::mutex> lk(d.mu);  // the workers are gone: no run is waiting
      }
      d.cv.notify_all();
      d.th.join();
    }
    if (d.ipb) rp_ipb_free(d.ipb);
    if (d.stream) (void)hipStreamDestroy(d.stream);
    for (int h = 0; h < kIpAux; h++) {
      if (d.aux[h]) (void)hipStreamDestroy(d.aux[h]);
      if (d.asm_done[h]) (void)hipEventDestroy(d.asm_done[h]);
    }
    if (d.solved) (void)hipEventDestroy(d.solved);
  }
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& j : s->jobs)
    if (j->ctx) rp_destroy(j->ctx);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->h_phase) (void)hipHostFree(s->h_phase);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}
