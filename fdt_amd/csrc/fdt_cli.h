// fdt_cli.h -- the xfibres option surface and the acquisition files, shared
// by the xfibres_gpu driver and the in-process bedpostx_gpu pipeline.
//
// Options follow xfibresoptions.h:82-148 with FSL Option<T> syntax
// ("--name=value", "--name value", "-x value"); bvals / bvecs are read as
// xfibres_gpu.cc:41-62 reads them (either orientation, bvecs normalised).
#pragma once
#include <sys/stat.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "fdt_xfibres.h"

namespace fdt_cli {

struct Cli {
  bool verbose = false, help = false, forcedir = false;
  std::string logdir = "logdir", data, mask, bvecs, bvals, gradnonlin;
  bool data_set = false, mask_set = false, bvecs_set = false, bvals_set = false, grad_set = false;
  bool savedyads = false;
  fdt_opts o{};
  int burnin_noard = 0;
  bool localinit = false;
};

inline void usage() {
  std::cerr << "Part of FSL-compatible bedpostx_gpu (MI355X engine)\n"
               "xfibres_gpu [options] <subject_dir> <part_id> <num_parts> <total_num_voxels>\n\n"
               "Compulsory arguments (You MUST set one or more of):\n"
               "\t-k,--data,--datafile\tdata file\n"
               "\t-m,--mask,--maskfile\tmask file\n"
               "\t-r,--bvecs\tb vectors file\n"
               "\t-b,--bvals\tb values file\n\n"
               "Optional arguments (You may optionally specify one or more of):\n"
               "\t-V,--verbose\tswitch on diagnostic messages\n"
               "\t-h,--help\tdisplay this message\n"
               "\t--ld,--logdir\tlog directory (default is logdir)\n"
               "\t--forcedir\tUse the actual directory name given - i.e. don't add + to make a new directory\n"
               "\t--nf,--nfibres\tMaximum number of fibres to fit in each voxel (default 1)\n"
               "\t--model\tWhich model to use. 1=deconv. with sticks (default). 2=deconv. with sticks and a range of diffusivities. 3=deconv. with zeppelins\n"
               "\t--fudge\tARD fudge factor\n"
               "\t--nj,--njumps\tNum of jumps to be made by MCMC (default is 1250)\n"
               "\t--bi,--burnin\tTotal num of jumps at start of MCMC to be discarded (default is 1000)\n"
               "\t--bn,--burnin_noard\tnum of burnin jumps before the ard is imposed (default is 0)\n"
               "\t--se,--sampleevery\tNum of jumps for each sample (MCMC) (default is 25)\n"
               "\t--upe,--updateproposalevery\tNum of jumps for each update to the proposal density std (MCMC) (default is 40)\n"
               "\t--seed\tseed for pseudo random number generator\n"
               "\t--noard\tTurn ARD off on all fibres\n"
               "\t--allard\tTurn ARD on on all fibres\n"
               "\t--nospat\tInitialise with tensor, not spatially\n"
               "\t--nonlinear\tInitialise with nonlinear fitting\n"
               "\t--cnonlinear\tInitialise with constrained nonlinear fitting\n"
               "\t--rician\tUse Rician noise modelling\n"
               "\t--f0\tAdd to the model an unattenuated signal compartment\n"
               "\t--ardf0\tUse ard on f0\n"
               "\t--Rmean\tSet the prior mean for R of model 3 (default:0.13- Must be<0.5)\n"
               "\t--Rstd\tSet the prior standard deviation for R of model 3 (default:0.03)\n"
               "\t--Rfudge\tIf set(>0), an ARD prior is used for R with the requested fudge factor\n"
               "\t--gradnonlin\tGradient Nonlinearity Tensor file\n"
               "\t--savedyads\twrite dyads<k>J, mean_f<k>samplesJ and dispersion<k>J as well (MI355X engine extension)\n"
               "\t--cpuinit\tStart the chains as the CPU xfibres does: without --nonlinear, --cnonlinear or --rician,\n"
               "\t\tfrom the tensor (--nospat) or the best finished neighbour (MI355X engine extension)\n";
}

// FSL Option<T> parsing: "--name=value", "--name value", "-x value"; flags take no value.
inline bool parse(int argc, char** argv, Cli& c, std::string& err) {
  fdt_default_opts(&c.o);
  struct Spec { const char* names; int kind; };  // kind: 0 flag, 1 string, 2 int, 3 float
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.empty() || a[0] != '-') { err = "unexpected argument: " + a; return false; }
    std::string name = a, val;
    bool has_eq = false;
    size_t eq = a.find('=');
    if (a.rfind("--", 0) == 0 && eq != std::string::npos) { name = a.substr(0, eq); val = a.substr(eq + 1); has_eq = true; }
    auto need = [&](std::string& out) -> bool {
      if (has_eq) { out = val; return true; }
      if (i + 1 >= argc) { err = "option " + name + " requires an argument"; return false; }
      out = argv[++i];
      return true;
    };
    auto as_int = [&](int& dst) -> bool {
      std::string s; if (!need(s)) return false;
      char* e = nullptr; long v = std::strtol(s.c_str(), &e, 10);
      if (!e || *e) { err = "bad integer for " + name + ": " + s; return false; }
      dst = (int)v; return true;
    };
    auto as_u64 = [&](uint64_t& dst) -> bool {
      std::string s; if (!need(s)) return false;
      char* e = nullptr; long long v = std::strtoll(s.c_str(), &e, 10);
      if (!e || *e) { err = "bad integer for " + name + ": " + s; return false; }
      dst = (uint64_t)v; return true;
    };
    auto as_float = [&](float& dst) -> bool {
      std::string s; if (!need(s)) return false;
      char* e = nullptr; float v = std::strtof(s.c_str(), &e);
      if (!e || *e) { err = "bad number for " + name + ": " + s; return false; }
      dst = v; return true;
    };
    auto is = [&](std::initializer_list<const char*> l) { for (auto* n : l) if (name == n) return true; return false; };
    if (is({"-V", "--verbose"})) c.verbose = true;
    else if (is({"-h", "--help"})) c.help = true;
    else if (is({"--ld", "--logdir"})) { if (!need(c.logdir)) return false; }
    else if (is({"--forcedir"})) c.forcedir = true;
    else if (is({"-k", "--data", "--datafile"})) { if (!need(c.data)) return false; c.data_set = true; }
    else if (is({"-m", "--mask", "--maskfile"})) { if (!need(c.mask)) return false; c.mask_set = true; }
    else if (is({"-r", "--bvecs"})) { if (!need(c.bvecs)) return false; c.bvecs_set = true; }
    else if (is({"-b", "--bvals"})) { if (!need(c.bvals)) return false; c.bvals_set = true; }
    else if (is({"--nf", "--nfibres"})) { if (!as_int(c.o.nfibres)) return false; }
    else if (is({"--model"})) { if (!as_int(c.o.model)) return false; }
    else if (is({"--fudge"})) { if (!as_float(c.o.fudge)) return false; }
    else if (is({"--nj", "--njumps"})) { if (!as_int(c.o.njumps)) return false; }
    else if (is({"--bi", "--burnin"})) { if (!as_int(c.o.nburn)) return false; }
    else if (is({"--bn", "--burnin_noard"})) { if (!as_int(c.burnin_noard)) return false; }
    else if (is({"--se", "--sampleevery"})) { if (!as_int(c.o.sampleevery)) return false; }
    else if (is({"--upe", "--updateproposalevery"})) { if (!as_int(c.o.updateproposalevery)) return false; }
    else if (is({"--seed"})) { if (!as_u64(c.o.seed)) return false; }
    else if (is({"--noard"})) c.o.no_ard = 1;
    else if (is({"--allard"})) c.o.all_ard = 1;
    else if (is({"--nospat"})) { c.localinit = true; c.o.nospat = 1; }
    else if (is({"--cpuinit"})) c.o.cpu_init = 1;
    else if (is({"--nonlinear"})) c.o.nonlin = 1;
    else if (is({"--cnonlinear"})) c.o.cnonlin = 1;
    else if (is({"--rician"})) c.o.rician = 1;
    else if (is({"--f0"})) c.o.f0 = 1;
    else if (is({"--ardf0"})) c.o.ardf0 = 1;
    else if (is({"--Rmean"})) { if (!as_float(c.o.R_prior_mean)) return false; }
    else if (is({"--Rstd"})) { if (!as_float(c.o.R_prior_std)) return false; }
    else if (is({"--Rfudge"})) { if (!as_float(c.o.R_prior_fudge)) return false; }
    else if (is({"--gradnonlin"})) { if (!need(c.gradnonlin)) return false; c.grad_set = true; }
    else if (is({"--savedyads"})) c.savedyads = true;
    else { err = "unknown option: " + name; return false; }
  }
  return true;
}

inline bool is_dir(const std::string& p) { struct stat sb; return stat(p.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode); }

// read_ascii_matrix: whitespace-separated rows
inline bool read_ascii(const std::string& path, std::vector<std::vector<double>>& rows) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::vector<double> r;
    double v;
    while (ss >> v) r.push_back(v);
    if (!r.empty()) rows.push_back(r);
  }
  return !rows.empty();
}


// bvals [N] and bvecs 3 x N (row-major) as float, bvecs normalised to unit
// length (xfibres_gpu.cc:41-62): a column of bvals is read as a row, N x 3
// bvecs are transposed.
inline bool load_acquisition(const std::string& bvals_path, const std::string& bvecs_path, std::vector<float>& fbvals,
                             std::vector<float>& fbvecs, std::string& err) {
  std::vector<std::vector<double>> bvr, bcr;
  if (!read_ascii(bvals_path, bvr)) { err = "cannot read bvals file " + bvals_path; return false; }
  if (!read_ascii(bvecs_path, bcr)) { err = "cannot read bvecs file " + bvecs_path; return false; }
  std::vector<double> bvals;
  if (bvr.size() > 1) for (auto& r : bvr) bvals.push_back(r[0]);  // column -> row
  else bvals = bvr[0];
  std::vector<std::vector<double>> bv;  // 3 x N
  if (bcr.size() > 3) {
    bv.assign(3, std::vector<double>(bcr.size()));
    for (size_t i = 0; i < bcr.size(); ++i)
      for (int k = 0; k < 3; ++k) bv[k][i] = k < (int)bcr[i].size() ? bcr[i][k] : 0.0;
  } else bv = bcr;
  int N = (int)bvals.size();
  if (N <= 0) { err = "The number of gradient directions must be greater than 0"; return false; }
  if (bv.size() != 3 || (int)bv[0].size() != N || (int)bv[1].size() != N || (int)bv[2].size() != N) {
    err = "xfibres_gpu error: The number of elements in bvals and number of vectors in bvecs must be the same\n";
    return false;
  }
  fbvals.assign(N, 0.0f);
  fbvecs.assign(3 * (size_t)N, 0.0f);
  for (int i = 0; i < N; ++i) {
    double s = std::sqrt(bv[0][i] * bv[0][i] + bv[1][i] * bv[1][i] + bv[2][i] * bv[2][i]);
    for (int k = 0; k < 3; ++k) fbvecs[k * N + i] = (float)(s != 0 ? bv[k][i] / s : bv[k][i]);
    fbvals[i] = (float)bvals[i];
  }
  return true;
}

}  // namespace fdt_cli
