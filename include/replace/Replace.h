// Replace.h -- drop-in replacement of the reference's public placer class.
//
// Same class name, namespace and public methods with the same signatures as
// the reference's include/replace/Replace.h:23-94.  The placement itself runs
// on the GPU through the C ABI of rp_capi.h.
//
// One additive member: setNetlist().  The reference reads its netlist from
// OpenDB (PlacerBase::init, src/placerBase.cpp:605-758); OpenDB is not part of
// this build, so the caller hands over the same data as a host SoA netlist.
// setDb()/setSta() are kept for signature compatibility and stored as
// borrowed pointers, as the reference does (src/replace.cpp:67-72).
#ifndef __REPLACE_HEADER__
#define __REPLACE_HEADER__

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "replace/rp_capi.h"

namespace odb {
class dbDatabase;
}
namespace sta {
class dbSta;
}

namespace replace {

class Replace {
 public:
  Replace();
  ~Replace();

  void init();
  void reset();

  void setDb(odb::dbDatabase* odb);
  void setSta(sta::dbSta* dbSta);

  void doInitialPlace();
  void doNesterovPlace();

  // Initial Place param settings
  void setInitialPlaceMaxIter(int iter);
  void setInitialPlaceMinDiffLength(int length);
  void setInitialPlaceMaxSolverIter(int iter);
  void setInitialPlaceMaxFanout(int fanout);
  void setInitialPlaceNetWeightScale(float scale);

  void setNesterovPlaceMaxIter(int iter);

  void setBinGridCntX(int binGridCntX);
  void setBinGridCntY(int binGridCntY);

  void setTargetDensity(float density);
  void setTargetOverflow(float overflow);
  void setInitDensityPenalityFactor(float penaltyFactor);
  void setInitWireLengthCoef(float coef);
  void setMinPhiCoef(float minPhiCoef);
  void setMaxPhiCoef(float maxPhiCoef);

  // HPWL: half-parameter wire length.
  void setReferenceHpwl(float deltaHpwl);

  void setIncrementalPlaceMode(bool mode);
  void setVerboseLevel(int verbose);

  // ---- additive (no reference counterpart) ----
  // Host SoA netlist that stands in for OpenDB; borrowed until the next
  // doInitialPlace()/doNesterovPlace() returns.
  void setNetlist(const rp_netlist* netlist);
  // Names of the netlist's pins (by pin index; only top-level ports are
  // read), for the "[WARN] <port> toplevel port is not placed!" lines of
  // PlacerBase (src/placerBase.cpp:394-399).  Without names a port is
  // called "port<k>".
  void setPinNames(const std::vector<std::string>& names);
  // Instance lower-left corners after placement (what updateDb writes).
  bool getPositions(std::vector<int32_t>& lx, std::vector<int32_t>& ly) const;
  // When false, fatal REPL codes are returned by lastCode() instead of
  // calling exit(code) (the reference always exits, src/logger.cpp:31-36).
  void setExitOnError(bool on);
  // Poisson transform: 0 FP32 FFT (default), 1 split-BF16 GEMMs on matrix cores.
  void setPoissonTransform(int mode);
  int lastCode() const;
  const rp_nesterov_stats& lastStats() const;

 private:
  odb::dbDatabase* db_;
  sta::dbSta* sta_;
  const rp_netlist* netlist_;
  rp_ctx* ctx_;

  int initialPlaceMaxIter_;
  int initialPlaceMinDiffLength_;
  int initialPlaceMaxSolverIter_;
  int initialPlaceMaxFanout_;
  float initialPlaceNetWeightScale_;

  int nesterovPlaceMaxIter_;
  int binGridCntX_;
  int binGridCntY_;
  float overflow_;
  float density_;
  float initDensityPenalityFactor_;
  float initWireLengthCoef_;
  float minPhiCoef_;
  float maxPhiCoef_;
  float referenceHpwl_;

  bool incrementalPlaceMode_;
  int verbose_;

  bool exitOnError_;
  int poissonTransform_;
  int lastCode_;
  rp_nesterov_stats lastStats_;

  std::vector<std::string> pinNames_;

  rp_params params() const;
  bool ensureContext();
  void fatal(int code, const char* msg);
  void printPlacerInfo();
  void printNesterovBaseInfo(const rp_nb_info& nb);
  void printNesterovRun(int code, const std::vector<rp_iter_record>& rec);
};
}  // namespace replace

#endif
