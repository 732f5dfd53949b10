// replace_cli.cpp -- command-line driver with the flags of the reference's
// `global_placement` Tcl command (src/replace.tcl:1-97), running the
// replace::Replace facade on a netlist file written by
// replace_amd.netlist.Netlist.write_bin.
//
//   replace_cli NETLIST.bin [-density d] [-init_density_penalty p]
//       [-init_wirelength_coef c] [-min_phi_coef m] [-max_phi_coef m]
//       [-overflow o] [-initial_place_max_iter n] [-initial_place_max_fanout n]
//       [-bin_grid_count n] [-verbose_level n] [-skip_initial_place]
//       [-incremental] [-timing_driven] [-wire_res r] [-wire_cap c]
//       [-nesterov_max_iter n] [-poisson_mfma] [-out POSITIONS.txt]
//
// As in the Tcl procedure: the density defaults to 0.7, numeric values must
// be positive numbers (sta::check_positive_float / check_positive_integer:
// a number >= 0; the OpenSTA helpers are not part of the reference, so
// their message text is this driver's own), -skip_initial_place sets 0
// initial-place iterations, -incremental also sets the incremental mode,
// -bin_grid_count sets both axes, -timing_driven / -wire_res / -wire_cap are
// accepted and unused (the reference reads and ignores them), and a design
// without rows is refused.  -nesterov_max_iter, -poisson_mfma and -out are
// additions of this driver.  The exit code is the facade's (REPL codes).
#include <cerrno>
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
  std::vector<std::string> names;
  rp_netlist nl;
};

bool read_block(std::ifstream& f, std::vector<int32_t>& v, size_t n) {
  v.resize(n);
  if (n == 0) return true;
  return (bool)f.read(reinterpret_cast<char*>(v.data()), n * sizeof(int32_t));
}

// Header (17 int32): magic, version 2, dbu, die (4), site (2), rows, insts,
// nets, pins, pin boxes, inst pin starts, inst pins, name bytes; then the
// blocks in that order and the NUL-terminated pin names.
bool load(const char* path, NetlistFile& F) {
  std::ifstream f(path, std::ios::binary);
  if (!f || !read_block(f, F.head, 17) || F.head[0] != 0x52504E4C || F.head[1] != 2) return false;
  int rows = F.head[9], ni = F.head[10], nn = F.head[11], np = F.head[12];
  int nbox = F.head[13], nips = F.head[14], nipn = F.head[15], nnames = F.head[16];
  if (rows < 0 || ni < 0 || nn < 0 || np < 0 || nbox < 0 || nips < 0 || nipn < 0 || nnames < 0) return false;
  if (!read_block(f, F.rows, 4 * (size_t)rows) || !read_block(f, F.insts, 4 * (size_t)ni) ||
      !read_block(f, F.fixed32, ni) || !read_block(f, F.nstart, nn + 1) || !read_block(f, F.pinst, np) ||
      !read_block(f, F.poff, 2 * (size_t)np) || !read_block(f, F.pbox, 4 * (size_t)nbox) ||
      !read_block(f, F.ips, nips) || !read_block(f, F.ipn, nipn))
    return false;
  std::string blob(nnames, '\0');
  if (nnames && !f.read(&blob[0], nnames)) return false;
  for (size_t a = 0; a < blob.size();) {
    size_t b = blob.find('\0', a);
    if (b == std::string::npos) b = blob.size();
    F.names.push_back(blob.substr(a, b - a));
    a = b + 1;
  }
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

[[noreturn]] void arg_error(const std::string& flag, const char* value, const char* what) {
  printf("Error: %s '%s' is not a positive %s.\n", flag.c_str(), value, what);
  fflush(stdout);
  exit(1);
}

// sta::check_positive_float: a number, >= 0
float positive_float(const std::string& flag, const char* v) {
  char* end = nullptr;
  errno = 0;
  const double x = strtod(v, &end);
  if (!*v || *end || errno || !(x >= 0.0)) arg_error(flag, v, "float");
  return (float)x;
}

// sta::check_positive_integer: an integer, >= 0
int positive_int(const std::string& flag, const char* v) {
  char* end = nullptr;
  errno = 0;
  const long x = strtol(v, &end, 10);
  if (!*v || *end || errno || x < 0 || x > 2147483647L) arg_error(flag, v, "integer");
  return (int)x;
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
  rp.setPinNames(F.names);
  // global_placement: density defaults to 0.7 (src/replace.tcl:17-22)
  rp.setTargetDensity(0.7f);
  const char* out = nullptr;
  bool skip_ip = false, incremental = false;
  int ip_iter = -1;
  for (int i = 2; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&](void) -> const char* {
      if (i + 1 >= argc) {
        fprintf(stderr, "%s needs a value\n", a.c_str());
        exit(2);
      }
      return argv[++i];
    };
    if (a == "-density") rp.setTargetDensity(positive_float(a, next()));
    else if (a == "-init_density_penalty") rp.setInitDensityPenalityFactor(positive_float(a, next()));
    else if (a == "-init_wirelength_coef") rp.setInitWireLengthCoef(positive_float(a, next()));
    else if (a == "-min_phi_coef") rp.setMinPhiCoef(positive_float(a, next()));
    else if (a == "-max_phi_coef") rp.setMaxPhiCoef(positive_float(a, next()));
    else if (a == "-overflow") rp.setTargetOverflow(positive_float(a, next()));
    else if (a == "-initial_place_max_iter") ip_iter = positive_int(a, next());
    else if (a == "-initial_place_max_fanout") rp.setInitialPlaceMaxFanout(positive_int(a, next()));
    else if (a == "-verbose_level") rp.setVerboseLevel(positive_int(a, next()));
    else if (a == "-nesterov_max_iter") rp.setNesterovPlaceMaxIter(positive_int(a, next()));
    else if (a == "-bin_grid_count") {
      int n = positive_int(a, next());
      rp.setBinGridCntX(n);
      rp.setBinGridCntY(n);
    } else if (a == "-skip_initial_place") skip_ip = true;
    else if (a == "-incremental") incremental = true;
    else if (a == "-timing_driven") {
    }  // accepted, unused by the reference
    else if (a == "-wire_res" || a == "-wire_cap") (void)next();  // accepted, unused by the reference
    else if (a == "-out") out = next();
    else if (a == "-poisson_mfma") rp.setPoissonTransform(1);
    else {
      fprintf(stderr, "unknown flag %s\n", a.c_str());
      return 2;
    }
  }
  // src/replace.tcl:24-35: -skip_initial_place wins over -initial_place_max_iter;
  // -incremental sets 0 iterations and the incremental mode
  if (skip_ip) rp.setInitialPlaceMaxIter(0);
  else if (ip_iter >= 0) rp.setInitialPlaceMaxIter(ip_iter);
  if (incremental) {
    rp.setInitialPlaceMaxIter(0);
    rp.setIncrementalPlaceMode(true);
  }
  if (F.nl.num_rows <= 0) {  // src/replace.tcl:89-96
    printf("Error: no rows defined in design. Use initialize_floorplan to add rows.\n");
    return 0;
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
  fflush(stdout);
  return rp.lastCode();
}
