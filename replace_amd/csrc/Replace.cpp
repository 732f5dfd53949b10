// Replace.cpp -- the replace::Replace facade over the C ABI.
//
// Parameter defaults and reset() values follow the reference's
// src/replace.cpp:13-65; doInitialPlace / doNesterovPlace follow
// src/replace.cpp:73-128 and print the reference's log: the "[INFO]" blocks
// of PlacerBase and NesterovBase, the unplaced-port "[WARN]" lines, the
// "[InitialPlace]" and "[NesterovSolve]" lines, the verbose-3 detail of
// NesterovPlace (from the device's per-evaluation trace records), and the
// "[ERROR] msg (REPL-code)" + exit(code) of the fatal paths
// (src/logger.cpp:31-36), REPL-0 for a NaN divergence included.
#include "replace/Replace.h"

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

extern "C" int rp_get_ip_records(rp_ctx* c, float* err, int64_t* hpwl, int32_t cap, int32_t* count);

namespace replace {

Replace::Replace()
    : db_(nullptr),
      sta_(nullptr),
      netlist_(nullptr),
      ctx_(nullptr),
      initialPlaceMaxIter_(20),
      initialPlaceMinDiffLength_(1500),
      initialPlaceMaxSolverIter_(100),
      initialPlaceMaxFanout_(200),
      initialPlaceNetWeightScale_(800),
      nesterovPlaceMaxIter_(2000),
      binGridCntX_(0),
      binGridCntY_(0),
      overflow_(0.1f),
      density_(1.0f),
      initDensityPenalityFactor_(0.00008f),
      initWireLengthCoef_(0.25f),
      minPhiCoef_(0.95f),
      maxPhiCoef_(1.05f),
      referenceHpwl_(446000000),
      incrementalPlaceMode_(false),
      verbose_(0),
      exitOnError_(true),
      poissonTransform_(0),
      lastCode_(0) {
  memset(&lastStats_, 0, sizeof lastStats_);
}

Replace::~Replace() {
  reset();
  if (ctx_) rp_destroy(ctx_);
  ctx_ = nullptr;
}

void Replace::init() {}

void Replace::reset() {
  // the reference's reset() values (src/replace.cpp:42-65), used by its destructor
  db_ = nullptr;
  sta_ = nullptr;
  initialPlaceMaxIter_ = 20;
  initialPlaceMinDiffLength_ = 1500;
  initialPlaceMaxSolverIter_ = 100;
  initialPlaceMaxFanout_ = 200;
  initialPlaceNetWeightScale_ = 800;
  nesterovPlaceMaxIter_ = 2000;
  binGridCntX_ = binGridCntY_ = 0;
  overflow_ = 0;
  density_ = 0;
  initDensityPenalityFactor_ = 0.0001f;
  initWireLengthCoef_ = 1.0f;
  minPhiCoef_ = 0.95f;
  maxPhiCoef_ = 1.05f;
  referenceHpwl_ = 446000000;
  incrementalPlaceMode_ = false;
  verbose_ = 0;
}

void Replace::setDb(odb::dbDatabase* db) { db_ = db; }
void Replace::setSta(sta::dbSta* sta) { sta_ = sta; }
void Replace::setNetlist(const rp_netlist* netlist) { netlist_ = netlist; }
void Replace::setExitOnError(bool on) { exitOnError_ = on; }
void Replace::setPoissonTransform(int mode) { poissonTransform_ = mode; }
int Replace::lastCode() const { return lastCode_; }
const rp_nesterov_stats& Replace::lastStats() const { return lastStats_; }

rp_params Replace::params() const {
  rp_params p;
  rp_params_default(&p);
  p.initial_place_max_iter = initialPlaceMaxIter_;
  p.initial_place_min_diff_length = initialPlaceMinDiffLength_;
  p.initial_place_max_solver_iter = initialPlaceMaxSolverIter_;
  p.initial_place_max_fanout = initialPlaceMaxFanout_;
  p.initial_place_net_weight_scale = initialPlaceNetWeightScale_;
  p.nesterov_place_max_iter = nesterovPlaceMaxIter_;
  p.bin_grid_cnt_x = binGridCntX_;
  p.bin_grid_cnt_y = binGridCntY_;
  p.target_overflow = overflow_;
  p.target_density = density_;
  p.init_density_penalty = initDensityPenalityFactor_;
  p.init_wire_length_coef = initWireLengthCoef_;
  p.min_phi_coef = minPhiCoef_;
  p.max_phi_coef = maxPhiCoef_;
  p.reference_hpwl = referenceHpwl_;
  p.incremental_place_mode = incrementalPlaceMode_ ? 1 : 0;
  p.verbose = verbose_;
  p.poisson_transform = poissonTransform_;
  return p;
}

// Log lines in the reference's formats (src/logger.cpp:39-90): printf and
// iostream output share stdout in order (iostreams are synchronised with
// stdio by default).
namespace {
void info_int(const char* k, long long v) { printf("[INFO] %s = %lld\n", k, v); }
void info_pair(const char* k, long long a, long long b) { printf("[INFO] %s = (%lld, %lld)\n", k, a, b); }
void info_float(const char* k, float v) { printf("[INFO] %s = %.6f\n", k, v); }
void info_sig(const char* k, float v) { printf("[INFO] %s = %g\n", k, v); }
}  // namespace

void Replace::fatal(int code, const char* msg) {
  lastCode_ = code;
  printf("[ERROR] %s (REPL-%d)\n", msg, code);
  fflush(stdout);
  if (exitOnError_) exit(code);
}

void Replace::setPinNames(const std::vector<std::string>& names) { pinNames_ = names; }

// PlacerBase::init and printInfo (src/placerBase.cpp:609-626, 394-399, 902-950)
void Replace::printPlacerInfo() {
  rp_pb_info pb;
  if (rp_get_pb_info(ctx_, &pb) != 0) return;
  info_int("DBU", pb.dbu_per_micron);
  info_pair("SiteSize", pb.site_w, pb.site_h);
  info_pair("CoreAreaLxLy", pb.core_lx, pb.core_ly);
  info_pair("CoreAreaUxUy", pb.core_ux, pb.core_uy);
  if (pb.num_unplaced_ports > 0 && netlist_) {
    // pins are net-major, so ports come in the order PlacerBase creates them
    for (int q = 0; q < netlist_->num_pins; q++) {
      if (netlist_->pin_inst[q] >= 0) continue;
      const int32_t* b = netlist_->pin_box ? netlist_->pin_box + 4 * (size_t)q : nullptr;
      const bool unplaced = !b || b[0] == INT32_MAX || b[1] == INT32_MAX || b[2] == INT32_MIN || b[3] == INT32_MIN;
      if (!unplaced) continue;
      std::string name = (size_t)q < pinNames_.size() && !pinNames_[q].empty() ? pinNames_[q]
                                                                              : "port" + std::to_string(q);
      printf("[WARN] %s toplevel port is not placed!\n       Replace will regard %s is placed in (0, 0) (REPL-1)\n",
             name.c_str(), name.c_str());
    }
  }
  info_int("NumInstances", pb.num_insts);
  info_int("NumPlaceInstances", pb.num_place_insts);
  info_int("NumFixedInstances", pb.num_fixed_insts);
  info_int("NumDummyInstances", pb.num_dummy_insts);
  info_int("NumNets", pb.num_nets);
  info_int("NumPins", pb.num_pins);
  info_pair("DieAreaLxLy", pb.die_lx, pb.die_ly);
  info_pair("DieAreaUxUy", pb.die_ux, pb.die_uy);
  info_pair("CoreAreaLxLy", pb.core_lx, pb.core_ly);
  info_pair("CoreAreaUxUy", pb.core_ux, pb.core_uy);
  info_int("CoreArea", pb.core_area);
  info_int("NonPlaceInstsArea", pb.nonplace_insts_area);
  info_int("PlaceInstsArea", pb.place_insts_area);
  info_float("Util(%)", pb.util);
  info_int("StdInstsArea", pb.std_insts_area);
  info_int("MacroInstsArea", pb.macro_insts_area);
  fflush(stdout);
}

bool Replace::ensureContext() {
  if (ctx_) return true;
  if (!netlist_) {
    fatal(1, "no netlist: call setNetlist() before placing");
    return false;
  }
  rp_params p = params();
  int r = rp_create(netlist_, &p, &ctx_);
  if (r < 0) {
    std::string m = ctx_ ? rp_last_error(ctx_) : "rp_create failed";
    fatal(r, m.c_str());
    return false;
  }
  printPlacerInfo();
  if (r == RP_ERR_FILLER) {
    // PlacerBase::printInfo (src/placerBase.cpp:952-956)
    printf("Error: Util exceeds 100%%.\n       Please double-check your die/row size\n");
    fflush(stdout);
    lastCode_ = 1;
    if (exitOnError_) exit(1);
    return false;
  }
  return true;
}

void Replace::doInitialPlace() {
  // a fresh PlacerBase per call (src/replace.cpp:73-88)
  if (ctx_) rp_destroy(ctx_);
  ctx_ = nullptr;
  lastCode_ = 0;
  if (!ensureContext()) return;
  if (verbose_ >= 3) printf("[PROC] Begin InitialPlace ...\n");
  rp_ip_stats st;
  int r = rp_initial_place(ctx_, &st);
  if (r != 0) {
    fatal(r, rp_last_error(ctx_));
    return;
  }
  int32_t n = 0;
  rp_get_ip_records(ctx_, nullptr, nullptr, 0, &n);
  std::vector<float> err(n);
  std::vector<int64_t> hp(n);
  rp_get_ip_records(ctx_, err.data(), hp.data(), n, &n);
  for (int i = 0; i < n; i++)
    printf("[InitialPlace]  Iter: %d CG Error: %g HPWL: %lld\n", i + 1, err[i], (long long)hp[i]);
  if (verbose_ >= 3) printf("[PROC] End InitialPlace\n");
  fflush(stdout);
}

// NesterovBase::init (src/nesterovBase.cpp:645-713, 1003-1005, 1127-1133)
void Replace::printNesterovBaseInfo(const rp_nb_info& nb) {
  if (verbose_ >= 3) {
    info_int("FillerInit: CoreArea", nb.core_area);
    info_int("FillerInit: WhiteSpaceArea", nb.white_space_area);
    info_int("FillerInit: MovableArea", nb.movable_area);
    info_int("FillerInit: TotalFillerArea", nb.total_filler_area);
    info_int("FillerInit: NumFillerCells", nb.num_fillers);
    info_int("FillerInit: FillerCellArea", (long long)nb.filler_dx * nb.filler_dy);
    info_pair("FillerInit: FillerCellSize", nb.filler_dx, nb.filler_dy);
  }
  info_int("FillerInit: NumGCells", nb.num_gcells);
  info_int("FillerInit: NumGNets", nb.num_gnets);
  info_int("FillerInit: NumGPins", nb.num_gpins);
  info_float("TargetDensity", nb.target_density);
  info_int("AveragePlaceInstArea", nb.average_place_inst_area);
  info_int("IdealBinArea", nb.ideal_bin_area);
  info_int("IdealBinCnt", nb.ideal_bin_count);
  info_int("TotalBinArea", nb.total_bin_area);
  info_pair("BinCnt", nb.bin_cnt_x, nb.bin_cnt_y);
  info_pair("BinSize", nb.bin_size_x, nb.bin_size_y);
  info_sig("NumBins", (float)((long long)nb.bin_cnt_x * nb.bin_cnt_y));
  fflush(stdout);
}

// The log of NesterovPlace::init and doNesterovPlace (src/nesterovPlace.cpp:
// 64-175, 206-273, 315-471, 514-583): the "[NesterovSolve]" lines always,
// the verbose-3 detail from the per-evaluation trace records.
void Replace::printNesterovRun(int code, const std::vector<rp_iter_record>& rec) {
  std::vector<rp_eval_record> ev;
  if (verbose_ >= 3) {
    int32_t n = 0;
    rp_get_eval_records(ctx_, nullptr, 0, &n);
    ev.resize(n);
    rp_get_eval_records(ctx_, ev.data(), n, &n);
    ev.resize(n);
  }
  const bool v3 = verbose_ >= 3 && ev.size() >= 1;
  auto grads = [&](const rp_eval_record& e) {
    info_sig("  DensityPenalty", e.density_penalty);
    info_sig("  WireLengthGradSum", e.wl_grad_sum);
    info_sig("  DensityGradSum", e.density_grad_sum);
    info_sig("  GradSum", e.grad_sum);
  };
  size_t at = 0;  // next evaluation record
  float initStep = 0.0f, prevOverflow = 0.0f;
  int64_t prevHpwl = 0;
  if (v3) {
    rp_nesterov_state st;
    memset(&st, 0, sizeof st);
    rp_nesterov_get_state(ctx_, &st);
    const rp_eval_record& a = ev[0];
    printf("[PROC] Begin NesterovInit ...\n");
    info_int("InitialHPWL", a.hpwl);
    info_sig("BaseWireLengthCoef", st.base_wl_coef);
    info_sig("InitSumOverflow", a.overflow);
    info_sig("WireLengthCoef", a.wl_coef);
    grads(a);
    at = 1;
    prevHpwl = a.hpwl;
    if (!a.diverged && ev.size() >= 2) {
      const rp_eval_record& b = ev[1];
      grads(b);
      at = 2;
      if (!b.diverged) {
        info_sig("WireLengthGradSum", b.wl_grad_sum);
        info_sig("DensityGradSum", b.density_grad_sum);
        info_sig("InitDensityPenalty", (b.wl_grad_sum / b.density_grad_sum) * initDensityPenalityFactor_);
        info_sig("PrevSumOverflow", b.overflow);
        info_sig("  CoordinateDistance", b.coordinate_distance);
        info_sig("  GradientDistance", b.gradient_distance);
        info_sig("InitialStepLength", b.new_step_length);
        printf("[PROC] End NesterovInit\n");
        initStep = b.new_step_length;
        prevOverflow = b.overflow;
      }
    }
  }
  if (code == RP_ERR_INIT_DIVERGED) return;
  float curA = 1.0f, maxPhi = maxPhiCoef_;
  bool phiChanged = false;
  const int nIter = (int)rec.size();
  const int last = nIter + ((lastStats_.diverged && (code == 0 || code == RP_ERR_BACKTRACK)) ? 1 : 0);
  for (int i = 0; i < last; i++) {
    if (v3) {
      const float prevA = curA;
      curA = (float)((1.0 + sqrt(4.0 * (double)prevA * (double)prevA + 1.0)) * 0.5);
      const float coeff = (float)(((double)prevA - 1.0) / (double)curA);
      info_int("Iter", i + 1);
      info_sig("  PreviousA", prevA);
      info_sig("  CurrentA", curA);
      info_sig("  Coefficient", coeff);
      info_sig("  StepLength", i == 0 ? initStep : rec[i - 1].step_length);
      int trials = 0;
      const rp_eval_record* accepted = nullptr;
      while (at < ev.size() && ev[at].iter == i) {
        const rp_eval_record& e = ev[at++];
        trials++;
        grads(e);
        if (e.diverged) break;
        info_sig("  CoordinateDistance", e.coordinate_distance);
        info_sig("  GradientDistance", e.gradient_distance);
        info_sig("  NewStepLength", e.new_step_length);
        accepted = &e;
      }
      // numBackTrak + 1 (src/nesterovPlace.cpp:390): trials, or the cap + 1
      // when every trial was rejected
      const bool exhausted = i >= nIter && code == RP_ERR_BACKTRACK;
      info_int("  NumBackTrak", exhausted ? trials + 1 : trials);
      if (!phiChanged && prevOverflow < 0.35f) {
        phiChanged = true;
        maxPhi = (float)((double)maxPhi * 0.99);
      }
      if (i >= nIter) break;  // diverged in this iteration: no updateNextIter
      const rp_iter_record& r = rec[i];
      info_sig("  Gradient", accepted ? accepted->gradient_norm : 0.0f);
      info_sig("  Phi", accepted ? accepted->sum_phi : 0.0f);
      info_sig("  Overflow", r.overflow);
      info_int("  PreviousHPWL", prevHpwl);
      info_int("  NewHPWL", r.hpwl);
      const float scaled = (float)(r.hpwl - prevHpwl) / referenceHpwl_;
      info_sig("  InputScaleDiffHPWL", scaled);
      float coef = scaled < 0 ? maxPhi : (float)((double)maxPhi * pow((double)maxPhi, (double)scaled * -1.0));
      coef = minPhiCoef_ < coef ? coef : minPhiCoef_;
      info_sig("  PhiCoef", coef);
      prevHpwl = r.hpwl;
      prevOverflow = r.overflow;
    }
    if (i < nIter && (i == 0 || (i + 1) % 10 == 0)) {
      std::cout << "[NesterovSolve] Iter: " << i + 1 << " overflow: " << rec[i].overflow
                << " HPWL: " << rec[i].hpwl << std::endl;
    }
  }
  if (lastStats_.finished)
    std::cout << "[NesterovSolve] Finished with Overflow: " << lastStats_.final_overflow << std::endl;
  fflush(stdout);
}

void Replace::doNesterovPlace() {
  lastCode_ = 0;
  if (!ensureContext()) return;
  rp_nb_info nb;
  int r = rp_nesterov_setup(ctx_, &nb);
  if (r < 0) {
    fatal(r, rp_last_error(ctx_));
    return;
  }
  if (r == RP_ERR_FILLER) {
    // NesterovBase::initFillerGCells (src/nesterovBase.cpp:1116-1121)
    fatal(1,
          "Filler area is negative!!\n       Please put higher target density or \n       Re-floorplan to have "
          "enough coreArea\n");
    return;
  }
  printNesterovBaseInfo(nb);
  r = rp_nesterov_place(ctx_, &lastStats_);
  if (r < 0) {
    fatal(r, rp_last_error(ctx_));
    return;
  }
  int32_t n = 0;
  rp_get_iter_records(ctx_, nullptr, 0, &n);
  std::vector<rp_iter_record> rec(n);
  rp_get_iter_records(ctx_, rec.data(), n, &n);
  printNesterovRun(r, rec);
  if (r == RP_ERR_INIT_DIVERGED) {
    fatal(2, "RePlAce diverged. Please tune the parameters again");
    return;
  }
  if (r == RP_ERR_BACKTRACK)
    fatal(3, "RePlAce divergence detected. \n        Please decrease init_density_penalty value");
  else if (r == RP_ERR_REBOUND)
    fatal(4, "RePlAce divergence detected. \n        Please decrease max_phi_cof value");
  else if (lastStats_.diverged)
    fatal(0, "");  // a NaN / inf gradient sum (src/nesterovPlace.cpp:371-374, 466-471)
}

bool Replace::getPositions(std::vector<int32_t>& lx, std::vector<int32_t>& ly) const {
  if (!ctx_ || !netlist_) return false;
  lx.resize(netlist_->num_insts);
  ly.resize(netlist_->num_insts);
  return rp_get_positions(ctx_, lx.data(), ly.data()) == 0;
}

void Replace::setInitialPlaceMaxIter(int iter) { initialPlaceMaxIter_ = iter; }
void Replace::setInitialPlaceMinDiffLength(int length) { initialPlaceMinDiffLength_ = length; }
void Replace::setInitialPlaceMaxSolverIter(int iter) { initialPlaceMaxSolverIter_ = iter; }
void Replace::setInitialPlaceMaxFanout(int fanout) { initialPlaceMaxFanout_ = fanout; }
void Replace::setInitialPlaceNetWeightScale(float scale) { initialPlaceNetWeightScale_ = scale; }
void Replace::setNesterovPlaceMaxIter(int iter) { nesterovPlaceMaxIter_ = iter; }
void Replace::setBinGridCntX(int binGridCntX) { binGridCntX_ = binGridCntX; }
void Replace::setBinGridCntY(int binGridCntY) { binGridCntY_ = binGridCntY; }
void Replace::setTargetOverflow(float overflow) { overflow_ = overflow; }
void Replace::setTargetDensity(float density) { density_ = density; }
void Replace::setInitDensityPenalityFactor(float penaltyFactor) { initDensityPenalityFactor_ = penaltyFactor; }
void Replace::setInitWireLengthCoef(float coef) { initWireLengthCoef_ = coef; }
void Replace::setMinPhiCoef(float minPhiCoef) { minPhiCoef_ = minPhiCoef; }
void Replace::setMaxPhiCoef(float maxPhiCoef) { maxPhiCoef_ = maxPhiCoef; }
void Replace::setReferenceHpwl(float refHpwl) { referenceHpwl_ = refHpwl; }
void Replace::setIncrementalPlaceMode(bool mode) { incrementalPlaceMode_ = mode; }
void Replace::setVerboseLevel(int verbose) { verbose_ = verbose; }

}  // namespace replace
