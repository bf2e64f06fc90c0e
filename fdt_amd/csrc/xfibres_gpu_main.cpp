// xfibres_gpu -- drop-in host driver for bedpostx_gpu on MI355X.
//
// Same command line and files as the reference driver xfibres_gpu.cc:27-243:
//   xfibres_gpu --data=<S>.bedpostX/data_<p> --mask=<mask> -b <bvals> -r <bvecs>
//               --forcedir --logdir=<dir> [xfibres options] <S> <p> <NJOBS> <nvox>
// The options are the xfibres surface (xfibresoptions.h:82-148); the last four
// arguments are positional (xfibres_gpu.cc:25, :69-99).  The part file is a
// raw matrix of `Real` [N][nvox_part] (volume-major, xfibres_gpu.cc:104-113);
// outputs are raw `Real` [nsamples][nvox_part] files "<name>J" in --logdir
// (xfibres_gpu.cc:194-222, CUDA/samples.cu:262-347), read by merge_parts_gpu.
// `Real` is newmat's element type, double in FSL builds; FDT_REAL_BYTES=4 in
// the environment switches to float files.
//
// Differences by design: the part is processed in one resident pass (288 GB
// of HBM hold any part; the reference splits into 12,800-voxel sub-parts for
// old GPUs, CUDA/options/options.h:1-3), so no sub-part files are joined; and
// the chains are seeded by --seed and the global voxel index
// (part * (nvox / NJOBS) + voxel), so results do not depend on NJOBS.
#include <sys/stat.h>
#include <sys/time.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "fdt_xfibres.h"

namespace {

struct Cli {
  bool verbose = false, help = false, forcedir = false;
  std::string logdir = "logdir", data, mask, bvecs, bvals, gradnonlin;
  bool data_set = false, mask_set = false, bvecs_set = false, bvals_set = false, grad_set = false;
  bool savedyads = false;
  fdt_opts o{};
  int burnin_noard = 0;
  bool localinit = false;
};

void usage() {
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
               "\t--savedyads\twrite dyads<k>J, mean_f<k>samplesJ and dispersion<k>J as well (MI355X engine extension)\n";
}

// FSL Option<T> parsing: "--name=value", "--name value", "-x value"; flags take no value.
bool parse(int argc, char** argv, Cli& c, std::string& err) {
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
    else if (is({"--nospat"})) c.localinit = true;
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

bool is_dir(const std::string& p) { struct stat sb; return stat(p.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode); }

// FSL Log::makeDir / setthenmakeDir: without --forcedir a name that exists
// gets '+' appended until it is new.
std::string make_logdir(const std::string& want, bool force) {
  std::string d = want;
  if (!force) while (is_dir(d)) d += "+";
  if (!is_dir(d)) mkdir(d.c_str(), 0777);
  return d;
}

// read_ascii_matrix: whitespace-separated rows
bool read_ascii(const std::string& path, std::vector<std::vector<double>>& rows) {
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

bool write_part(const std::string& path, const float* src, size_t n, int real_bytes) {
  std::ofstream out(path, std::ios::binary);
  if (!out) return false;
  if (real_bytes == 4) out.write(reinterpret_cast<const char*>(src), n * sizeof(float));
  else {
    std::vector<double> d(src, src + n);
    out.write(reinterpret_cast<const char*>(d.data()), n * sizeof(double));
  }
  return bool(out);
}

void write_monitor(const std::string& subjdir, int n) {
  std::string f = subjdir + ".bedpostX/logs/monitor/" + std::to_string(n);
  std::ofstream out(f, std::ios::binary);
  if (out) out.write("done", 4);
}

}  // namespace

int main(int argc, char* argv[]) {
  auto t0 = std::chrono::steady_clock::now();
  const char* usage_tail =
      "xfibres_gpu. The last 4 arguments must be:\n\tSubject-directory \n\tNumThisPart \n\tTotalNumParts \n\tTotalNumVoxels(all parts)";
  if (argc < 5) { usage(); std::cerr << usage_tail << std::endl; return EXIT_FAILURE; }
  Cli c;
  std::string err;
  if (!parse(argc - 4, argv, c, err)) { usage(); std::cerr << "\n" << err << std::endl; return 2; }
  if (c.help || !c.data_set || !c.mask_set || !c.bvecs_set || !c.bvals_set) { usage(); return 2; }
  std::string logdir = make_logdir(c.logdir, c.forcedir);
  std::cout << "Log directory is: " << logdir << std::endl;

  int real_bytes = 8;
  if (const char* rb = std::getenv("FDT_REAL_BYTES")) real_bytes = std::atoi(rb) == 4 ? 4 : 8;

  // bvals / bvecs (xfibres_gpu.cc:41-62)
  std::vector<std::vector<double>> bvr, bcr;
  if (!read_ascii(c.bvals, bvr)) { std::cerr << "cannot read bvals file " << c.bvals << std::endl; return EXIT_FAILURE; }
  if (!read_ascii(c.bvecs, bcr)) { std::cerr << "cannot read bvecs file " << c.bvecs << std::endl; return EXIT_FAILURE; }
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
  if (N <= 0) { std::cerr << "The number of gradient directions must be greater than 0" << std::endl; return EXIT_FAILURE; }
  if (bv.size() != 3 || (int)bv[0].size() != N || (int)bv[1].size() != N || (int)bv[2].size() != N) {
    std::cerr << "xfibres_gpu error: The number of elements in bvals and number of vectors in bvecs must be the same\n" << std::endl;
    return EXIT_FAILURE;
  }
  std::vector<float> fbvals(N), fbvecs(3 * N);
  for (int i = 0; i < N; ++i) {
    double s = std::sqrt(bv[0][i] * bv[0][i] + bv[1][i] * bv[1][i] + bv[2][i] * bv[2][i]);
    for (int k = 0; k < 3; ++k) fbvecs[k * N + i] = (float)(s != 0 ? bv[k][i] / s : bv[k][i]);
    fbvals[i] = (float)bvals[i];
  }

  // positional arguments (xfibres_gpu.cc:69-99)
  std::string subjdir = argv[argc - 4];
  if (!is_dir(subjdir)) {
    std::cerr << usage_tail << std::endl << "\nSubject-directory: " << subjdir << " is not a directory" << std::endl;
    return EXIT_FAILURE;
  }
  auto parse_int = [](const char* s, int& v) { char* e = nullptr; long x = std::strtol(s, &e, 10); v = (int)x; return e && *e == 0 && e != s; };
  int idPart, nParts, totalNvox;
  if (!parse_int(argv[argc - 3], idPart)) {
    std::cerr << usage_tail << std::endl << "\nNumThisPart: " << argv[argc - 3] << " is not a valid number" << std::endl;
    return EXIT_FAILURE;
  }
  if (!parse_int(argv[argc - 2], nParts) || nParts <= 0) {
    std::cerr << usage_tail << std::endl << "\nTotalNumParts: " << argv[argc - 2] << " is not a valid number" << std::endl;
    return EXIT_FAILURE;
  }
  if (!parse_int(argv[argc - 1], totalNvox) || totalNvox <= 0) {
    std::cerr << usage_tail << std::endl << "\nThe number of voxels must be greater than 0" << std::endl;
    return EXIT_FAILURE;
  }
  if (idPart < 0 || idPart >= nParts) {
    std::cerr << usage_tail << std::endl << "\nNumThisPart: " << idPart << " is outside 0.." << nParts - 1 << std::endl;
    return EXIT_FAILURE;
  }
  // this part of the data (xfibres_gpu.cc:104-113)
  int size_part = totalNvox / nParts;
  uint64_t vox_offset = (uint64_t)idPart * (uint64_t)size_part;
  if (idPart == nParts - 1) size_part = totalNvox - size_part * (nParts - 1);
  // raw `Real` [rows][size_part] files: the data part and the grad_dev part
  // (xfibres_gpu.cc:108-122; the gradient deviations are 9 rows)
  auto read_real = [&](const std::string& path, const char* what, int rows, std::vector<float>& out) {
    out.assign((size_t)rows * size_part, 0.0f);
    std::ifstream in(path, std::ios::binary);
    if (!in) { std::cerr << "cannot open " << what << " file " << path << std::endl; return false; }
    if (real_bytes == 4) in.read(reinterpret_cast<char*>(out.data()), out.size() * sizeof(float));
    else {
      std::vector<double> d(out.size());
      in.read(reinterpret_cast<char*>(d.data()), d.size() * sizeof(double));
      for (size_t i = 0; i < d.size(); ++i) out[i] = (float)d[i];
    }
    if (!in) { std::cerr << what << " file " << path << " holds fewer than " << rows << " x " << size_part << " values" << std::endl; return false; }
    return true;
  };
  std::vector<float> data, grad;
  if (!read_real(c.data, "data", N, data)) return EXIT_FAILURE;
  if (c.grad_set && !read_real(c.gradnonlin, "gradnonlin", 9, grad)) return EXIT_FAILURE;
  std::cout << "Number of Voxels to compute in this part: " << size_part << std::endl;
  std::cout << "Number of Directions: " << N << std::endl;
  if (c.o.rician && !c.o.nonlin)
    std::cout << "Rician noise model requested. Non-linear parameter initialization will be performed, overriding other initialization options!" << std::endl << std::endl;

  fdt_ctx* ctx = fdt_create(0, &c.o);
  if (!ctx) { std::cerr << "xfibres_gpu: cannot create the engine (options or device)" << std::endl; return EXIT_FAILURE; }
  auto fail = [&](const char* what) {
    std::cerr << "xfibres_gpu: " << what << ": " << fdt_last_error(ctx) << std::endl;
    fdt_destroy(ctx);
    return EXIT_FAILURE;
  };
  std::cout << "SubPart 1 of 1: processing " << size_part << " voxels" << std::endl;
  if (fdt_set_acquisition(ctx, N, fbvals.data(), fbvecs.data())) return fail("acquisition");
  if (fdt_load_data(ctx, data.data(), FDT_LAYOUT_VOL_MAJOR, size_part)) return fail("data");
  if (c.grad_set && fdt_load_grad_dev(ctx, grad.data(), FDT_LAYOUT_VOL_MAJOR, size_part)) return fail("gradnonlin");
  if (fdt_run(ctx, vox_offset) || fdt_sync(ctx)) return fail("run");

  const int K = c.o.nfibres, ns = c.o.njumps / c.o.sampleevery;
  size_t nsv = (size_t)ns * size_part;
  std::vector<float> th((size_t)K * nsv), ph((size_t)K * nsv), f((size_t)K * nsv);
  std::vector<float> mS0(size_part), md(size_part), mdstd(size_part), mR(size_part), mf0(size_part), mtau(size_part);
  std::vector<float> mf((size_t)K * size_part), dy((size_t)K * 3 * size_part), disp((size_t)K * size_part);
  fdt_outputs out{};
  out.th = th.data(); out.ph = ph.data(); out.f = f.data();
  out.mean_S0 = mS0.data(); out.mean_d = md.data(); out.mean_dstd = mdstd.data(); out.mean_R = mR.data();
  out.mean_f0 = mf0.data(); out.mean_tau = mtau.data(); out.mean_f = mf.data(); out.dyads = dy.data(); out.disp = disp.data();
  if (fdt_get_outputs(ctx, &out)) return fail("outputs");
  fdt_timing tm{};
  fdt_get_timing(ctx, &tm);
  fdt_destroy(ctx);

  // GPU phase log (CUDA/xfibres_gpu.cu:55-65 and the per-phase timers)
  {
    char part_s[16];
    std::snprintf(part_s, sizeof part_s, "%04d", idPart);
    std::ofstream lg(subjdir + ".bedpostX/logs/logs_gpu/part_" + part_s + "-subpart_0000", std::ios::app);
    if (lg) {
      lg << "----------------------------------------------------- \n";
      lg << "---------------- PART " << idPart << " SUBPART 0 ------------------- \n";
      lg << "----------------------------------------------------- \n\n";
      lg << "FIT TIME: " << tm.ms_fit / 1000.0 << " seconds\n";
      lg << "MCMC TIME (burn-in + record): " << tm.ms_mcmc / 1000.0 << " seconds\n";
      lg << "FINISH TIME: " << tm.ms_finish / 1000.0 << " seconds\n";
      lg << "MCMC tile: " << tm.mcmc_lanes << " lanes x " << tm.mcmc_vpl << " volumes per voxel\n";
    }
  }
  if (nParts == 1) write_monitor(subjdir, 0);

  // outputs (xfibres_gpu.cc:194-222)
  auto wr = [&](const std::string& name, const float* p, size_t n) {
    if (!write_part(logdir + "/" + name + "J", p, n, real_bytes)) {
      std::cerr << "xfibres_gpu: cannot write " << logdir << "/" << name << "J" << std::endl;
      std::exit(EXIT_FAILURE);
    }
  };
  wr("mean_dsamples", md.data(), size_part);
  if (c.o.model >= 2) wr("mean_d_stdsamples", mdstd.data(), size_part);
  if (c.o.model == 3) wr("mean_Rsamples", mR.data(), size_part);
  if (c.o.f0) wr("mean_f0samples", mf0.data(), size_part);
  if (c.o.rician) wr("mean_tausamples", mtau.data(), size_part);
  for (int k = 0; k < K; ++k) {
    wr("th" + std::to_string(k + 1) + "samples", th.data() + (size_t)k * nsv, nsv);
    wr("ph" + std::to_string(k + 1) + "samples", ph.data() + (size_t)k * nsv, nsv);
    wr("f" + std::to_string(k + 1) + "samples", f.data() + (size_t)k * nsv, nsv);
    if (c.savedyads) {
      wr("mean_f" + std::to_string(k + 1) + "samples", mf.data() + (size_t)k * size_part, size_part);
      wr("dyads" + std::to_string(k + 1), dy.data() + (size_t)k * 3 * size_part, (size_t)3 * size_part);
      wr("dispersion" + std::to_string(k + 1), disp.data() + (size_t)k * size_part, size_part);
    }
  }
  wr("mean_S0samples", mS0.data(), size_part);

  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << std::endl << "Part processed in: " << secs << " seconds" << std::endl;
  if (nParts > 1) write_monitor(subjdir, idPart);
  return 0;
}
