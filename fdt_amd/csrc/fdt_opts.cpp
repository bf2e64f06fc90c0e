// fdt_opts.cpp -- the xfibres option defaults (xfibresoptions.h:104-148), in
// host-only code so the option parser of the drivers builds without HIP (the
// sanitizer build of the host parsers links it).
#include <cstring>

#include "fdt_xfibres.h"

extern "C" void fdt_default_opts(fdt_opts* o) {
  std::memset(o, 0, sizeof *o);
  o->nfibres = 1;
  o->model = 1;
  o->fudge = 1.0f;
  o->njumps = 1250;
  o->nburn = 1000;
  o->sampleevery = 25;
  o->updateproposalevery = 40;
  o->lm_fp64 = 2;  // FP32 normal equations, FP64 parameters and cost
  o->seed = 8665904;
  o->R_prior_mean = 0.13f;
  o->R_prior_std = 0.03f;
  o->R_prior_fudge = 0.0f;
}
