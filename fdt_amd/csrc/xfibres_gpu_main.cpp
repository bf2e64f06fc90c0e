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

#include "fdt_cli.h"
#include "fdt_xfibres.h"

namespace {

using namespace fdt_cli;

// FSL Log::makeDir / setthenmakeDir: without --forcedir a name that exists
// gets '+' appended until it is new.
std::string make_logdir(const std::string& want, bool force) {
  std::string d = want;
  if (!force) while (is_dir(d)) d += "+";
  if (!is_dir(d)) mkdir(d.c_str(), 0777);
  return d;
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
  std::vector<float> fbvals, fbvecs;
  if (!load_acquisition(c.bvals, c.bvecs, fbvals, fbvecs, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
  const int N = (int)fbvals.size();

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

  if (c.o.cpu_init && !c.o.nospat && !c.o.rician && !c.o.nonlin && !c.o.cnonlin) {
    std::cerr << "xfibres_gpu: --cpuinit needs --nospat here: the part files carry no voxel grid for the neighbour "
                 "initialisation (bedpostx_gpu has it)" << std::endl;
    return EXIT_FAILURE;
  }
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
