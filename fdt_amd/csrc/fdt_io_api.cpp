// fdt_io_api.cpp -- C ABI of the NIfTI reader/writer (include/fdt_io.h).
#include "fdt_io.h"

#include <cstring>
#include <string>

#include "fdt_nifti.h"

struct fdt_image {
  fdt::Nifti n;
};

namespace {
void set_err(char* err, int errlen, const std::string& e) {
  if (err && errlen > 0) {
    std::strncpy(err, e.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = 0;
  }
}
}  // namespace

extern "C" {

fdt_image* fdt_image_read(const char* path, char* err, int errlen) {
  if (!path) { set_err(err, errlen, "null path"); return nullptr; }
  std::string p = fdt::image_path(path);
  if (p.empty()) { set_err(err, errlen, std::string("image not found: ") + path); return nullptr; }
  fdt_image* im = new fdt_image;
  std::string e;
  if (!fdt::read_nifti(p, im->n, e)) { set_err(err, errlen, e); delete im; return nullptr; }
  return im;
}

void fdt_image_free(fdt_image* im) { delete im; }

void fdt_image_dims(const fdt_image* im, int* dim, float* pixdim) {
  for (int i = 0; i < 8; ++i) {
    if (dim) dim[i] = im ? im->n.dim[i] : 0;
    if (pixdim) pixdim[i] = im ? im->n.pixdim[i] : 0.0f;
  }
}

const float* fdt_image_data(const fdt_image* im) { return im ? im->n.data.data() : nullptr; }

int fdt_image_write(const char* path, const fdt_image* like, int nt, const float* data, float cal_min, float cal_max,
                    char* err, int errlen) {
  if (!path || !like || !data || nt < 1) { set_err(err, errlen, "bad argument"); return -1; }
  std::string e;
  if (!fdt::write_nifti(path, like->n, nt, data, cal_min, cal_max, e)) { set_err(err, errlen, e); return -1; }
  return 0;
}

long fdt_image_mask_count(const fdt_image* mask) { return mask ? (long)fdt::mask_indices(mask->n).size() : -1; }

long fdt_image_mask_matrix(const fdt_image* img, const fdt_image* mask, float* out) {
  if (!img || !mask || !out) return -1;
  for (int i = 1; i <= 3; ++i)
    if (img->n.dim[i] != mask->n.dim[i]) return -1;
  std::vector<size_t> idx = fdt::mask_indices(mask->n);
  fdt::mask_matrix(img->n, idx, out);
  return (long)idx.size();
}

}  // extern "C"
