// xfibres_pred -- the posterior-predictive signal of a bedpostx result on
// MI355X (the command line and files of xfibres_pred.cc:16-198):
//
//   xfibres_pred bpx_dir [data_file] output
//
// bpx_dir holds bvals, bvecs, mean_dsamples, dyads<k>, mean_f<k>samples, and
// optionally mean_S0samples (else the mean of the b <= 50 volumes of
// data_file), mean_d_stdsamples (model 2), mean_Rsamples (model 3),
// mean_f0samples and nodif_brain_mask.  The prediction over the mask voxels
// runs on the GPU (fdt_predict); output is a 4D image with one volume per
// acquisition volume on the grid of mean_dsamples.
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "fdt_cli.h"
#include "fdt_nifti.h"
#include "fdt_xfibres.h"

int main(int argc, char* argv[]) {
  if (argc < 3 || argc > 4) {
    std::cerr << " \nusage: xfibres_pred bpx_dir [data_file] output\n\n"
                 "       bpx_dir is the bedpostx output directory \n"
                 "       data_file is optional, in case no mean_S0samples exists in the bpx dir, get the mean S0 from "
                 "the data \n \n";
    return EXIT_FAILURE;
  }
  const std::string dir = argv[1];
  std::string err;
  std::vector<float> bvals, bvecs;
  if (!fdt_cli::load_acquisition(dir + "/bvals", dir + "/bvecs", bvals, bvecs, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
  const int N = (int)bvals.size();
  auto exists = [&](const std::string& name) { return !fdt::image_path(dir + "/" + name).empty(); };
  auto load = [&](const std::string& name, fdt::Nifti& img) {
    if (!fdt::read_nifti(fdt::image_path(dir + "/" + name), img, err)) { std::cerr << err << std::endl; std::exit(EXIT_FAILURE); }
  };
  int K = 0;
  while (exists("dyads" + std::to_string(K + 1))) ++K;
  if (!exists("mean_dsamples")) { std::cout << "No mean_dsamples file exists!" << std::endl; return EXIT_FAILURE; }
  fdt::Nifti d, S0, dstd, R, f0, mask;
  load("mean_dsamples", d);
  const size_t nxyz = d.nxyz();
  auto same_grid = [&](const fdt::Nifti& a) { return a.dim[1] == d.dim[1] && a.dim[2] == d.dim[2] && a.dim[3] == d.dim[3]; };
  if (!exists("mean_S0samples") && argc != 4) { std::cerr << "No mean_S0samples or data file exists!" << std::endl; return EXIT_FAILURE; }
  if (!exists("mean_S0samples")) {
    const std::string data_name = fdt::image_path(argv[2]);
    if (data_name.empty()) { std::cerr << "Wrong data filename!" << std::endl; return EXIT_FAILURE; }
    std::cout << "No mean_SOsamples found, getting the average of b=0's from data..." << std::endl;
    fdt::Nifti data;
    if (!fdt::read_nifti(data_name, data, err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    if (!same_grid(data) || data.nt() != N) { std::cerr << "data does not match mean_dsamples and bvals" << std::endl; return EXIT_FAILURE; }
    // volumes with 0 <= b <= 50 count as b = 0 (xfibres_pred.cc:75-81): float sums, then / count
    S0 = d;
    S0.data.assign(nxyz, 0.0f);
    int b0count = 0;
    for (int l = 0; l < N; ++l)
      if (bvals[l] >= 0.0f && bvals[l] <= 50.0f) {
        for (size_t v = 0; v < nxyz; ++v) S0.data[v] += data.data[(size_t)l * nxyz + v];
        ++b0count;
      }
    for (auto& v : S0.data) v /= (float)b0count;
  } else {
    load("mean_S0samples", S0);
  }
  std::vector<fdt::Nifti> dy(K), fk(K);
  for (int n = 0; n < K; ++n) {
    load("dyads" + std::to_string(n + 1), dy[n]);
    if (!exists("mean_f" + std::to_string(n + 1) + "samples")) {
      std::cerr << "No mean_f" << n + 1 << "samples file exists!" << std::endl;
      return EXIT_FAILURE;
    }
    load("mean_f" + std::to_string(n + 1) + "samples", fk[n]);
  }
  int model = 1;
  if (exists("mean_d_stdsamples")) { model = 2; load("mean_d_stdsamples", dstd); }
  if (exists("mean_Rsamples")) { model = 3; load("mean_Rsamples", R); }
  const bool has_f0 = exists("mean_f0samples");
  if (has_f0) load("mean_f0samples", f0);
  std::cout << "Files for model" << model << " with " << K << " fibres found" << std::endl;
  if (has_f0) std::cout << "Also an f0 noise floor is assumed" << std::endl << std::endl;
  if (exists("nodif_brain_mask")) load("nodif_brain_mask", mask);
  else { mask = d; mask.data.assign(nxyz, 1.0f); }
  for (const fdt::Nifti* im : {&S0, &mask})
    if (!same_grid(*im)) { std::cerr << "images of " << dir << " are on different grids" << std::endl; return EXIT_FAILURE; }

  // the mask voxels as columns, the prediction on the GPU
  const std::vector<size_t> idx = fdt::mask_indices(mask);
  const int nvox = (int)idx.size();
  auto col = [&](const fdt::Nifti& im, int nt) {
    std::vector<float> m((size_t)nvox * nt);  // [t][nvox]
    for (int t = 0; t < nt; ++t)
      for (int i = 0; i < nvox; ++i) m[(size_t)t * nvox + i] = im.data[(size_t)t * nxyz + idx[i]];
    return m;
  };
  std::vector<float> cS0 = col(S0, 1), cd = col(d, 1), cdstd, cR, cf0, cf, cdy;
  if (model >= 2) cdstd = col(dstd, 1);
  if (model == 3) cR = col(R, 1);
  if (has_f0) cf0 = col(f0, 1);
  for (int n = 0; n < K; ++n) {
    std::vector<float> a = col(fk[n], 1), b = col(dy[n], 3);
    cf.insert(cf.end(), a.begin(), a.end());
    cdy.insert(cdy.end(), b.begin(), b.end());
  }
  std::vector<float> pred((size_t)N * nvox);
  int rc = fdt_predict(0, N, bvals.data(), bvecs.data(), nvox, K, model, cS0.data(), cd.data(),
                       model >= 2 ? cdstd.data() : nullptr, model == 3 ? cR.data() : nullptr,
                       has_f0 ? cf0.data() : nullptr, cf.data(), cdy.data(), pred.data());
  if (rc) { std::cerr << "xfibres_pred: the prediction failed (" << rc << ")" << std::endl; return EXIT_FAILURE; }
  std::vector<float> vol(nxyz * (size_t)N);
  fdt::scatter_volumes(idx, nxyz, N, pred.data(), vol.data());
  std::cout << "saving results" << std::endl;
  std::string out = argv[argc - 1];
  if (out.size() < 4 || (out.compare(out.size() - 4, 4, ".nii") != 0 && (out.size() < 7 || out.compare(out.size() - 7, 7, ".nii.gz") != 0))) {
    const char* ot = std::getenv("FSLOUTPUTTYPE");
    out += (ot && std::string(ot) == "NIFTI") ? ".nii" : ".nii.gz";
  }
  if (!fdt::write_nifti(out, d, N, vol.data(), 0.0f, 0.0f, err, 16)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
  return 0;
}
