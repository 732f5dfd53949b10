// Replace.cpp -- the replace::Replace facade over the C ABI.
//
// Parameter defaults and reset() values follow the reference's
// src/replace.cpp:13-65; doInitialPlace / doNesterovPlace follow
// src/replace.cpp:73-128 and print the same unconditional log lines
// ("[InitialPlace]  Iter: ...", "[NesterovSolve] Iter: ...",
// src/initialPlace.cpp:88-90, src/nesterovPlace.cpp:421-423, 460) and the
// same "[ERROR] msg (REPL-code)" + exit(code) on the fatal paths
// (src/logger.cpp:31-36).
#include "replace/Replace.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

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

void Replace::fatal(int code, const char* msg) {
  lastCode_ = code;
  std::cout << "[ERROR] " << msg << " (REPL-" << code << ")" << std::endl;
  if (exitOnError_) exit(code);
}

bool Replace::ensureContext() {
  if (ctx_) return true;
  if (!netlist_) {
    fatal(1, "no netlist: call setNetlist() before placing");
    return false;
  }
  rp_params p = params();
  int r = rp_create(netlist_, &p, &ctx_);
  if (r != 0) {
    if (r == RP_ERR_FILLER) {
      // PlacerBase::printInfo (src/placerBase.cpp:952-956)
      std::cout << "Error: Util exceeds 100%." << std::endl;
      std::cout << "       Please double-check your die/row size" << std::endl;
      lastCode_ = 1;
      if (exitOnError_) exit(1);
    } else {
      std::string m = ctx_ ? rp_last_error(ctx_) : "rp_create failed";
      fatal(r, m.c_str());
    }
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
    std::cout << "[InitialPlace]  Iter: " << i + 1 << " CG Error: " << err[i] << " HPWL: " << hp[i] << std::endl;
}

void Replace::doNesterovPlace() {
  lastCode_ = 0;
  if (!ensureContext()) return;
  int r = rp_nesterov_place(ctx_, &lastStats_);
  if (r < 0) {
    fatal(r, rp_last_error(ctx_));
    return;
  }
  if (r == RP_ERR_FILLER) {
    fatal(1, "Filler area is negative!!\n Please put higher target density or \n Re-floorplan to have enough coreArea\n");
    return;
  }
  int32_t n = 0;
  rp_get_iter_records(ctx_, nullptr, 0, &n);
  std::vector<rp_iter_record> rec(n);
  rp_get_iter_records(ctx_, rec.data(), n, &n);
  if (r == RP_ERR_INIT_DIVERGED) {
    fatal(2, "RePlAce diverged. Please tune the parameters again");
    return;
  }
  for (int i = 0; i < n; i++) {
    if (i == 0 || (i + 1) % 10 == 0)
      std::cout << "[NesterovSolve] Iter: " << i + 1 << " overflow: " << rec[i].overflow
                << " HPWL: " << rec[i].hpwl << std::endl;
  }
  if (lastStats_.finished)
    std::cout << "[NesterovSolve] Finished with Overflow: " << lastStats_.final_overflow << std::endl;
  if (r == RP_ERR_BACKTRACK)
    fatal(3, "RePlAce divergence detected. \n        Please decrease init_density_penalty value");
  else if (r == RP_ERR_REBOUND)
    fatal(4, "RePlAce divergence detected. \n        Please decrease max_phi_cof value");
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
