// bedpostx_gpu -- the whole bedpostx_gpu pipeline in one process on MI355X:
// NIfTI subject directory in, <subject>.bedpostX results out, no FSL tools.
//
// The reference runs a shell pipeline (CUDA/bedpostx_gpu): split_parts_gpu
// (CUDA/split_parts_gpu.cc:41-148: data.matrix(mask) cut into NJOBS raw part
// files), one xfibres_gpu process per part, then bedpostx_postproc_gpu.sh
// (:20-73): merge_parts_gpu (merge_parts_gpu.cc:17-172: parts -> merged_*
// and mean_* volumes), fslmaths -Tmean, make_dyadic_vectors
// (make_dyadic_vectors.cc:13-103), maskdyads (maskdyads:18-24), the
// dyads<k>_thr0.05_modf<k> and mean_fsumsamples images, xfms/eye.mat.
//
// Here the masked voxels go to the engine as one resident matrix (HBM holds a
// whole brain), the chains run in one launch, and the outputs are written
// straight from the engine's sorted samples, means and dyads: no part files,
// no merge.  Results equal those of the part-wise run for any NJOBS, because
// chains are seeded by --seed and the voxel's index in the mask.
//
//   bedpostx_gpu <subject_directory> [-n nf] [-w fudge] [-b burnin] [-j njumps]
//                [-s sampleevery] [-model 1|2|3] [-g] [xfibres options]
// (-NJOBS, -Q and -QSYS are accepted for compatibility and have no effect.)
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "fdt_cli.h"
#include "fdt_nifti.h"
#include "fdt_xfibres.h"

namespace {

using fdt::Nifti;
using Clock = std::chrono::steady_clock;

double secs(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }
bool file_exists(const std::string& p) { struct stat sb; return stat(p.c_str(), &sb) == 0 && S_ISREG(sb.st_mode); }

void usage() {
  std::cout << "\nUsage: bedpostx_gpu <subject_directory> [options]\n\n"
               "expects to find bvals and bvecs in subject directory\n"
               "expects to find data and nodif_brain_mask in subject directory\n"
               "expects to find grad_dev in subject directory, if -g is set\n\n"
               "<options>:\n"
               "-n (number of fibres per voxel, default 3)\n"
               "-w (ARD weight, more weight means less secondary fibres per voxel, default 1)\n"
               "-b (burnin period, default 1000)\n"
               "-j (number of jumps, default 1250)\n"
               "-s (sample every, default 25)\n"
               "-model (Deconvolution model. 1: with sticks, 2: with sticks with a range of diffusivities (default), "
               "3: with zeppelins)\n"
               "-g (consider gradient nonlinearities, default off)\n"
               "-NJOBS, -Q, -QSYS (accepted; the pipeline runs in this process on one GPU)\n\n"
               "ALTERNATIVELY: you can pass on xfibres options onto directly bedpostx_gpu\n"
               " For example:  bedpostx_gpu <subject directory> --noard --cnonlinear\n";
}

bool copy_file(const std::string& from, const std::string& to) {
  std::ifstream in(from, std::ios::binary);
  std::ofstream out(to, std::ios::binary);
  if (!in || !out) return false;
  out << in.rdbuf();
  return bool(out);
}

std::string image_ext(const std::string& path) {
  if (path.size() > 7 && path.compare(path.size() - 7, 7, ".nii.gz") == 0) return ".nii.gz";
  if (path.size() > 4 && path.compare(path.size() - 4, 4, ".nii") == 0) return ".nii";
  return "";
}

// one output image: nt volumes of the mask geometry
struct Out {
  std::string name;
  int nt;
  std::vector<float> vol;
  float cal_min, cal_max;
};

}  // namespace

int main(int argc, char* argv[]) {
  auto t_all = Clock::now();
  if (argc < 2 || std::string(argv[1]).empty()) { usage(); return EXIT_FAILURE; }
  // subject directory, absolute, no trailing '/'
  std::string subjdir = argv[1];
  if (fdt_cli::is_dir(subjdir)) {
    char buf[4096];
    std::string cwd = getcwd(buf, sizeof buf) ? buf : ".";
    if (subjdir[0] != '/') subjdir = cwd + "/" + subjdir;
  }
  while (subjdir.size() > 1 && subjdir.back() == '/') subjdir.pop_back();
  std::cout << "---------------------------------------------\n"
               "------------ BedpostX GPU Version -----------\n"
               "---------------------------------------------\n"
               "subjectdir is " << subjdir << std::endl;

  // bedpostx options (CUDA/bedpostx_gpu:130-158) -> xfibres options
  int nfibres = 3, burnin = 1000, njumps = 1250, sampleevery = 25, model = 2;
  std::string fudge = "1";
  bool gflag = false;
  std::vector<std::string> other;
  for (int i = 2; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&](std::string& dst) {
      if (i + 1 >= argc) { std::cerr << "option " << a << " requires an argument" << std::endl; std::exit(EXIT_FAILURE); }
      dst = argv[++i];
    };
    auto ival = [&](int& dst) {
      std::string s;
      val(s);
      char* e = nullptr;
      long v = std::strtol(s.c_str(), &e, 10);
      if (!e || *e) { std::cerr << "bad integer for " << a << ": " << s << std::endl; std::exit(EXIT_FAILURE); }
      dst = (int)v;
    };
    std::string ignored;
    if (a == "-QSYS" || a == "-Q" || a == "-NJOBS") val(ignored);
    else if (a == "-n") ival(nfibres);
    else if (a == "-w") val(fudge);
    else if (a == "-b") ival(burnin);
    else if (a == "-j") ival(njumps);
    else if (a == "-s") ival(sampleevery);
    else if (a == "-model") ival(model);
    else if (a == "-g") gflag = true;
    else other.push_back(a);
  }
  std::vector<std::string> xargs = {"bedpostx_gpu", "--nf=" + std::to_string(nfibres), "--fudge=" + fudge,
                                    "--bi=" + std::to_string(burnin), "--nj=" + std::to_string(njumps),
                                    "--se=" + std::to_string(sampleevery), "--model=" + std::to_string(model)};
  // --cpuinit (the CPU xfibres starts) replaces the script's constrained nonlinear fit; an explicit
  // --cnonlinear or --nonlinear among the options still selects the fit
  if (std::find(other.begin(), other.end(), "--cpuinit") == other.end()) xargs.push_back("--cnonlinear");
  xargs.insert(xargs.end(), other.begin(), other.end());
  std::vector<char*> xv;
  for (auto& s : xargs) xv.push_back(&s[0]);
  fdt_cli::Cli c;
  std::string err;
  if (!fdt_cli::parse((int)xv.size(), xv.data(), c, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
  if (c.help) { usage(); return EXIT_FAILURE; }

  // required files (CUDA/bedpostx_gpu:173-219)
  if (!fdt_cli::is_dir(subjdir)) { std::cerr << "subject directory " << argv[1] << " not found" << std::endl; return EXIT_FAILURE; }
  const std::string S = subjdir, B = subjdir + ".bedpostX";
  std::string bvecs_src = S + "/bvecs", bvals_src = S + "/bvals";
  if (!file_exists(bvecs_src)) {
    if (file_exists(S + "/bvecs.txt")) bvecs_src = S + "/bvecs.txt";
    else { std::cerr << S << "/bvecs not found" << std::endl; return EXIT_FAILURE; }
  }
  if (!file_exists(bvals_src)) {
    if (file_exists(S + "/bvals.txt")) bvals_src = S + "/bvals.txt";
    else { std::cerr << S << "/bvals not found" << std::endl; return EXIT_FAILURE; }
  }
  const std::string data_path = fdt::image_path(S + "/data");
  if (data_path.empty()) { std::cerr << S << "/data not found" << std::endl; return EXIT_FAILURE; }
  std::string grad_path;
  if (gflag) {
    grad_path = fdt::image_path(S + "/grad_dev");
    if (grad_path.empty()) { std::cerr << S << "/grad_dev not found" << std::endl; return EXIT_FAILURE; }
  }
  const std::string mask_path = fdt::image_path(S + "/nodif_brain_mask");
  if (mask_path.empty()) { std::cerr << S << "/nodif_brain_mask not found" << std::endl; return EXIT_FAILURE; }
  if (file_exists(B + "/xfms/eye.mat")) {
    std::cerr << S << " has already been processed: " << B << ".\n"
              << "Delete or rename " << B << " before repeating the process." << std::endl;
    return EXIT_FAILURE;
  }

  std::cout << "Making bedpostx directory structure" << std::endl;
  for (const std::string& d : {B, B + "/logs", B + "/logs/logs_gpu", B + "/logs/monitor", B + "/xfms"}) mkdir(d.c_str(), 0777);
  std::cout << "Copying files to bedpost directory" << std::endl;
  if (!copy_file(bvecs_src, B + "/bvecs") || !copy_file(bvals_src, B + "/bvals") ||
      !copy_file(mask_path, B + "/nodif_brain_mask" + image_ext(mask_path))) {
    std::cerr << "cannot copy the acquisition files and the mask into " << B << std::endl;
    return EXIT_FAILURE;
  }

  // inputs (split_parts_gpu.cc:96-119)
  auto t_in = Clock::now();
  std::vector<float> fbvals, fbvecs;
  if (!fdt_cli::load_acquisition(B + "/bvals", B + "/bvecs", fbvals, fbvecs, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
  const int N = (int)fbvals.size();
  Nifti mask, data, grad;
  if (!fdt::read_nifti(mask_path, mask, err) || !fdt::read_nifti(data_path, data, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
  auto same_grid = [&](const Nifti& a) { return a.dim[1] == mask.dim[1] && a.dim[2] == mask.dim[2] && a.dim[3] == mask.dim[3]; };
  if (!same_grid(data)) { std::cerr << "data and nodif_brain_mask have different dimensions" << std::endl; return EXIT_FAILURE; }
  if (data.nt() != N) {
    std::cerr << "split_parts_gpu error: The number of elements in bvals, number of vectors in bvecs and number of vols "
                 "in data must be the same\n" << std::endl;
    return EXIT_FAILURE;
  }
  if (gflag) {
    if (!fdt::read_nifti(grad_path, grad, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    if (!same_grid(grad) || grad.nt() != 9) { std::cerr << "grad_dev must hold 9 volumes on the mask's grid" << std::endl; return EXIT_FAILURE; }
  }
  if (const std::string nodif = fdt::image_path(S + "/nodif"); !nodif.empty()) {
    Nifti nd;
    if (fdt::read_nifti(nodif, nd, err) && same_grid(nd)) {  // fslmaths nodif -mas nodif_brain_mask nodif_brain
      std::vector<float> v(nd.data.begin(), nd.data.begin() + nd.nxyz());
      for (size_t i = 0; i < v.size(); ++i) if (mask.data[i] == 0.0f) v[i] = 0.0f;
      if (!fdt::write_nifti(B + "/nodif_brain.nii.gz", mask, 1, v.data(), 0.0f, 0.0f, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    }
  }
  const std::vector<size_t> idx = fdt::mask_indices(mask);
  const int nvox = (int)idx.size();
  if (nvox <= 0) { std::cerr << "The number of voxels and gradient directions must be greater than 0" << std::endl; return EXIT_FAILURE; }
  std::vector<float> dm((size_t)nvox * N), gm;
  fdt::mask_matrix(data, idx, dm.data());
  data.data.clear();
  data.data.shrink_to_fit();
  if (gflag) {
    gm.resize((size_t)nvox * 9);
    fdt::mask_matrix(grad, idx, gm.data());
  }
  const double s_in = secs(t_in);
  std::cout << "Pre-processing stage: " << nvox << " voxels, " << N << " volumes (" << s_in << " s)" << std::endl;

  // the engine: LM + MCMC + finishing over the whole mask
  auto t_run = Clock::now();
  fdt_ctx* ctx = fdt_create(0, &c.o);
  if (!ctx) { std::cerr << "bedpostx_gpu: cannot create the engine (options or device)" << std::endl; return EXIT_FAILURE; }
  auto fail = [&](const char* what) {
    std::cerr << "bedpostx_gpu: " << what << ": " << fdt_last_error(ctx) << std::endl;
    fdt_destroy(ctx);
    return EXIT_FAILURE;
  };
  if (fdt_set_acquisition(ctx, N, fbvals.data(), fbvecs.data())) return fail("acquisition");
  if (fdt_load_data(ctx, dm.data(), FDT_LAYOUT_VOX_MAJOR, nvox)) return fail("data");
  if (gflag && fdt_load_grad_dev(ctx, gm.data(), FDT_LAYOUT_VOX_MAJOR, nvox)) return fail("grad_dev");
  if (c.o.cpu_init) {  // the mask grid, for the neighbour initialisation of the CPU xfibres
    const std::vector<long long> gidx(idx.begin(), idx.end());
    if (fdt_set_grid(ctx, mask.dim[1], std::max(mask.dim[2], 1), std::max(mask.dim[3], 1), gidx.data())) return fail("grid");
  }
  if (fdt_run(ctx, 0) || fdt_sync(ctx)) return fail("run");
  const fdt_opts& o = c.o;
  const int K = o.nfibres, ns = o.njumps / o.sampleevery;
  const size_t nsv = (size_t)ns * nvox;
  std::vector<float> th((size_t)K * nsv), ph((size_t)K * nsv), f((size_t)K * nsv);
  std::vector<float> mS0(nvox), md(nvox), mdstd(nvox), mR(nvox), mf0(nvox), mtau(nvox);
  std::vector<float> mf((size_t)K * nvox), dy((size_t)K * 3 * nvox), disp((size_t)K * nvox);
  fdt_outputs out{};
  out.th = th.data(); out.ph = ph.data(); out.f = f.data();
  out.mean_S0 = mS0.data(); out.mean_d = md.data(); out.mean_dstd = mdstd.data(); out.mean_R = mR.data();
  out.mean_f0 = mf0.data(); out.mean_tau = mtau.data(); out.mean_f = mf.data(); out.dyads = dy.data(); out.disp = disp.data();
  if (fdt_get_outputs(ctx, &out)) return fail("outputs");
  fdt_timing tm{};
  fdt_get_timing(ctx, &tm);
  fdt_destroy(ctx);
  const double s_run = secs(t_run);

  // outputs (merge_parts_gpu.cc:133-166, CUDA/bedpostx_postproc_gpu.sh:20-50)
  auto t_out = Clock::now();
  const size_t nxyz = mask.nxyz();
  std::vector<Out> outs;
  auto vol_of = [&](const float* m, int nt) {
    std::vector<float> v(nxyz * (size_t)nt);
    fdt::scatter_volumes(idx, nxyz, nt, m, v.data());
    return v;
  };
  auto range = [](const std::vector<float>& v, float& mn, float& mx) {
    mn = v.empty() ? 0.0f : *std::min_element(v.begin(), v.end());
    mx = v.empty() ? 0.0f : *std::max_element(v.begin(), v.end());
  };
  // join_Parts display range: max / min of the volume where the reference passes -10
  auto add = [&](const std::string& name, const float* m, int nt, float cmin, float cmax, bool auto_min, bool auto_max) {
    Out r{name, nt, vol_of(m, nt), cmin, cmax};
    float mn, mx;
    range(r.vol, mn, mx);
    if (auto_min) r.cal_min = mn;
    if (auto_max) r.cal_max = mx;
    outs.push_back(std::move(r));
  };
  add("mean_dsamples", md.data(), 1, 0.0f, 0.0f, false, true);
  if (o.model >= 2) add("mean_d_stdsamples", mdstd.data(), 1, 0.0f, 0.0f, false, true);
  if (o.model == 3) add("mean_Rsamples", mR.data(), 1, 0.0f, 1.0f, false, false);
  if (o.f0) add("mean_f0samples", mf0.data(), 1, 0.0f, 1.0f, false, false);
  if (o.rician) add("mean_tausamples", mtau.data(), 1, 0.0f, 0.0f, false, true);
  add("mean_S0samples", mS0.data(), 1, 0.0f, 0.0f, false, true);
  // fslmaths -Tmean of the merged samples, per voxel in double
  auto tmean = [&](const float* m) {
    std::vector<float> r(nvox);
    for (int i = 0; i < nvox; ++i) {
      double s = 0.0;
      for (int t = 0; t < ns; ++t) s += m[(size_t)t * nvox + i];
      r[i] = (float)(s / ns);
    }
    return r;
  };
  std::vector<std::vector<float>> meanf(K);
  for (int k = 0; k < K; ++k) {
    const std::string n = std::to_string(k + 1);
    const float *tk = th.data() + (size_t)k * nsv, *pk = ph.data() + (size_t)k * nsv, *fk = f.data() + (size_t)k * nsv;
    add("merged_th" + n + "samples", tk, ns, 0.0f, 0.0f, true, true);
    add("merged_ph" + n + "samples", pk, ns, 0.0f, 0.0f, true, true);
    add("merged_f" + n + "samples", fk, ns, 0.0f, 1.0f, false, false);
    std::vector<float> mt = tmean(tk), mp = tmean(pk);
    meanf[k] = tmean(fk);
    add("mean_th" + n + "samples", mt.data(), 1, 0.0f, 0.0f, true, true);
    add("mean_ph" + n + "samples", mp.data(), 1, 0.0f, 0.0f, true, true);
    add("mean_f" + n + "samples", meanf[k].data(), 1, 0.0f, 1.0f, false, false);
    // make_dyadic_vectors: dyads<k> (3 volumes) and dyads<k>_dispersion, from the engine's finishing pass
    add("dyads" + n, dy.data() + (size_t)k * 3 * nvox, 3, -1.0f, 1.0f, false, false);
    add("dyads" + n + "_dispersion", disp.data() + (size_t)k * nvox, 1, 0.0f, 1.0f, false, false);
    if (k >= 1) {
      // maskdyads: mean_f<k> -thr 0.05 -bin -mul dyads<k>; then x (mean_f<k> / mean_f1), 0 where mean_f1 = 0
      std::vector<float> thr((size_t)3 * nvox), modf((size_t)3 * nvox);
      for (int i = 0; i < nvox; ++i) {
        const float keep = meanf[k][i] >= 0.05f ? 1.0f : 0.0f;
        const float ratio = meanf[0][i] != 0.0f ? meanf[k][i] / meanf[0][i] : 0.0f;
        for (int cpt = 0; cpt < 3; ++cpt) {
          const float v = keep * dy[((size_t)k * 3 + cpt) * nvox + i];
          thr[(size_t)cpt * nvox + i] = v;
          modf[(size_t)cpt * nvox + i] = v * ratio;
        }
      }
      add("dyads" + n + "_thr0.05", thr.data(), 3, -1.0f, 1.0f, false, false);
      add("dyads" + n + "_thr0.05_modf" + n, modf.data(), 3, -1.0f, 1.0f, false, false);
    }
  }
  {
    std::vector<float> fsum(nvox, 0.0f);
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < nvox; ++i) fsum[i] += meanf[k][i];
    add("mean_fsumsamples", fsum.data(), 1, 0.0f, 1.0f, false, false);
  }
  // FSLOUTPUTTYPE: NIFTI_GZ (FSL's default) or NIFTI; gzip level 1, members
  // compressed on all threads
  const char* ot = std::getenv("FSLOUTPUTTYPE");
  const std::string ext = (ot && std::string(ot) == "NIFTI") ? ".nii" : ".nii.gz";
  const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (auto& r : outs)
    if (!fdt::write_nifti(B + "/" + r.name + ext, mask, r.nt, r.vol.data(), r.cal_min, r.cal_max, err, nthreads)) {
      std::cerr << err << std::endl;
      return EXIT_FAILURE;
    }
  {
    std::ofstream eye(B + "/xfms/eye.mat");
    eye << "1 0 0 0\n0 1 0 0\n0 0 1 0\n0 0 0 1\n";
  }
  const double s_out = secs(t_out);
  {
    std::ofstream lg(B + "/logs/bedpostx_gpu.log", std::ios::app);
    lg << "voxels " << nvox << " volumes " << N << " fibres " << K << " model " << o.model << " samples " << ns << "\n"
       << "read + mask: " << s_in << " s\nengine (upload, LM, MCMC, finish, download): " << s_run << " s\n"
       << "  LM " << tm.ms_fit / 1000.0 << " s, MCMC " << tm.ms_mcmc / 1000.0 << " s, finish " << tm.ms_finish / 1000.0
       << " s (" << tm.mcmc_lanes << " lanes x " << tm.mcmc_vpl << " volumes per voxel)\n"
       << "write " << outs.size() << " images: " << s_out << " s\ntotal: " << secs(t_all) << " s\n";
  }
  std::cout << "Engine: " << s_run << " s (LM " << tm.ms_fit / 1000.0 << " s, MCMC " << tm.ms_mcmc / 1000.0
            << " s); wrote " << outs.size() << " images in " << s_out << " s" << std::endl;
  std::cout << "Done" << std::endl;
  return 0;
}
