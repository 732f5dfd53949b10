// replace_cli.cpp -- command-line driver with the flags of the reference's
// `global_placement` Tcl command (src/replace.tcl:8-97), running the
// replace::Replace facade on a netlist file written by
// replace_amd.netlist.Netlist.write_bin.
//
//   replace_cli NETLIST.bin [-density d] [-init_density_penalty p]
//       [-init_wirelength_coef c] [-min_phi_coef m] [-max_phi_coef m]
//       [-overflow o] [-initial_place_max_iter n] [-initial_place_max_fanout n]
//       [-bin_grid_count n] [-verbose_level n] [-skip_initial_place]
//       [-incremental] [-nesterov_max_iter n] [-poisson_mfma] [-out POSITIONS.txt]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "replace/Replace.h"

namespace {

struct NetlistFile {
  std::vector<int32_t> head, rows, insts, fixed32, nstart, pinst, poff, pbox, ips, ipn;
  std::vector<uint8_t> fixed;
  rp_netlist nl;
};

bool read_block(std::ifstream& f, std::vector<int32_t>& v, size_t n) {
  v.resize(n);
  if (n == 0) return true;
  return (bool)f.read(reinterpret_cast<char*>(v.data()), n * sizeof(int32_t));
}

bool load(const char* path, NetlistFile& F) {
  std::ifstream f(path, std::ios::binary);
  if (!f || !read_block(f, F.head, 17) || F.head[0] != 0x52504E4C) return false;
  int rows = F.head[9], ni = F.head[10], nn = F.head[11], np = F.head[12];
  int nbox = F.head[13], nips = F.head[14], nipn = F.head[15];
  if (!read_block(f, F.rows, 4 * (size_t)rows) || !read_block(f, F.insts, 4 * (size_t)ni) ||
      !read_block(f, F.fixed32, ni) || !read_block(f, F.nstart, nn + 1) || !read_block(f, F.pinst, np) ||
      !read_block(f, F.poff, 2 * (size_t)np) || !read_block(f, F.pbox, 4 * (size_t)nbox) ||
      !read_block(f, F.ips, nips) || !read_block(f, F.ipn, nipn))
    return false;
  F.fixed.assign(F.fixed32.begin(), F.fixed32.end());
  rp_netlist& nl = F.nl;
  nl.dbu_per_micron = F.head[2];
  nl.die_lx = F.head[3];
  nl.die_ly = F.head[4];
  nl.die_ux = F.head[5];
  nl.die_uy = F.head[6];
  nl.site_w = F.head[7];
  nl.site_h = F.head[8];
  nl.num_rows = rows;
  nl.row_box = F.rows.data();
  nl.num_insts = ni;
  nl.inst_box = F.insts.data();
  nl.inst_fixed = F.fixed.data();
  nl.num_nets = nn;
  nl.net_pin_start = F.nstart.data();
  nl.num_pins = np;
  nl.pin_inst = F.pinst.data();
  nl.pin_off = F.poff.data();
  nl.pin_box = nbox ? F.pbox.data() : nullptr;
  nl.inst_pin_start = nips ? F.ips.data() : nullptr;
  nl.inst_pins = nips ? F.ipn.data() : nullptr;
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s NETLIST.bin [global_placement flags] [-out FILE]\n", argv[0]);
    return 2;
  }
  NetlistFile F;
  if (!load(argv[1], F)) {
    fprintf(stderr, "cannot read netlist %s\n", argv[1]);
    return 2;
  }
  replace::Replace rp;
  rp.setNetlist(&F.nl);
  // global_placement: density defaults to 0.7 (src/replace.tcl:17-22)
  rp.setTargetDensity(0.7f);
  const char* out = nullptr;
  for (int i = 2; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&](void) -> const char* {
      if (i + 1 >= argc) {
        fprintf(stderr, "%s needs a value\n", a.c_str());
        exit(2);
      }
      return argv[++i];
    };
    if (a == "-density") rp.setTargetDensity((float)atof(next()));
    else if (a == "-init_density_penalty") rp.setInitDensityPenalityFactor((float)atof(next()));
    else if (a == "-init_wirelength_coef") rp.setInitWireLengthCoef((float)atof(next()));
    else if (a == "-min_phi_coef") rp.setMinPhiCoef((float)atof(next()));
    else if (a == "-max_phi_coef") rp.setMaxPhiCoef((float)atof(next()));
    else if (a == "-overflow") rp.setTargetOverflow((float)atof(next()));
    else if (a == "-initial_place_max_iter") rp.setInitialPlaceMaxIter(atoi(next()));
    else if (a == "-initial_place_max_fanout") rp.setInitialPlaceMaxFanout(atoi(next()));
    else if (a == "-verbose_level") rp.setVerboseLevel(atoi(next()));
    else if (a == "-nesterov_max_iter") rp.setNesterovPlaceMaxIter(atoi(next()));
    else if (a == "-bin_grid_count") {
      int n = atoi(next());
      rp.setBinGridCntX(n);
      rp.setBinGridCntY(n);
    } else if (a == "-skip_initial_place") rp.setInitialPlaceMaxIter(0);
    else if (a == "-incremental") {
      rp.setInitialPlaceMaxIter(0);
      rp.setIncrementalPlaceMode(true);
    } else if (a == "-out") out = next();
    else if (a == "-poisson_mfma") rp.setPoissonTransform(1);
    else {
      fprintf(stderr, "unknown flag %s\n", a.c_str());
      return 2;
    }
  }
  rp.doInitialPlace();
  rp.doNesterovPlace();
  if (out) {
    std::vector<int32_t> lx, ly;
    if (rp.getPositions(lx, ly)) {
      FILE* f = fopen(out, "w");
      if (f) {
        for (size_t k = 0; k < lx.size(); k++) fprintf(f, "%d %d\n", lx[k], ly[k]);
        fclose(f);
      }
    }
  }
  return 0;
}
