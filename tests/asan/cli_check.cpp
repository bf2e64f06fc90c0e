// cli_check.cpp -- drives the xfibres option parser and the bvals/bvecs
// reader of the drivers (fdt_amd/csrc/fdt_cli.h) on the command line it is
// given; built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/asan/Makefile, run by tests/test_asan_host.py.  Prints "OK ..." or
// "ERR <message>"; a sanitizer finding aborts with a report instead.
#include <cstdio>
#include <string>
#include <vector>

#include "fdt_cli.h"

int main(int argc, char** argv) {
  fdt_cli::Cli c;
  std::string err;
  if (!fdt_cli::parse(argc, argv, c, err)) {
    std::printf("ERR %s\n", err.c_str());
    return 0;
  }
  std::printf("OK nf=%d model=%d bi=%d nj=%d se=%d\n", c.o.nfibres, c.o.model, c.o.nburn, c.o.njumps, c.o.sampleevery);
  if (c.bvals_set && c.bvecs_set) {
    std::vector<float> bv, bc;
    if (fdt_cli::load_acquisition(c.bvals, c.bvecs, bv, bc, err)) std::printf("OK N=%zu\n", bv.size());
    else std::printf("ERR %s\n", err.c_str());
  }
  return 0;
}
