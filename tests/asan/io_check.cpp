// io_check.cpp -- drives the NIfTI-1 reader/writer and mask maps of the
// pipeline's host code (fdt_nifti.cpp through the include/fdt_io.h C ABI)
// over the image paths on the command line; built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/asan/Makefile, run by
// tests/test_asan_host.py.  Each path prints "OK <dims>" or "ERR <reason>";
// a sanitizer finding aborts with a report instead.
#include <cstdio>
#include <string>
#include <vector>

#include "fdt_io.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: io_check <scratch.nii> <image>...\n");
    return 2;
  }
  const char* scratch = argv[1];
  for (int a = 2; a < argc; ++a) {
    char err[256] = {0};
    fdt_image* im = fdt_image_read(argv[a], err, sizeof err);
    if (!im) {
      std::printf("ERR %s: %s\n", argv[a], err);
      continue;
    }
    int dim[8];
    float pix[8];
    fdt_image_dims(im, dim, pix);
    long n = 1;
    for (int i = 1; i <= dim[0] && i < 8; ++i) n *= dim[i] > 0 ? dim[i] : 1;
    const float* d = fdt_image_data(im);
    double s = 0;
    for (long i = 0; i < n; ++i) s += d[i];
    long nm = fdt_image_mask_count(im);
    long nt = dim[0] >= 4 && dim[4] > 0 ? dim[4] : 1;
    std::vector<float> m((size_t)(nm > 0 ? nm : 1) * nt);
    long got = fdt_image_mask_matrix(im, im, m.data());
    int rc = fdt_image_write(scratch, im, (int)nt, d, 0.0f, 1.0f, err, sizeof err);
    std::printf("OK %s: rank %d, %ld values, sum %g, mask %ld/%ld, write %d\n", argv[a], dim[0], n, s, nm, got, rc);
    fdt_image_free(im);
  }
  return 0;
}
