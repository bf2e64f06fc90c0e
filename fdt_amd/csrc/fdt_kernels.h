// fdt_kernels.h -- kernel argument blocks shared by the host orchestration and
// the HIP kernels (plain data only).
#pragma once
#include <stdint.h>

namespace fdt {

enum { KIND_SINGLE = 0, KIND_SINGLE_C = 1, KIND_MULTI = 2 };
enum { FIT_PIPELINE = 0, FIT_STAGE = 1, FIT_START = 2 };

constexpr int LM_VM = 8;       // volumes per lane in the LM kernel: N <= 512
constexpr int NPMAX = 14;      // LM parameters for K <= 3
constexpr int MAX_N = 512;     // acquisition length supported
constexpr int FIT_WAVES = 4;   // waves (voxels) per LM workgroup

struct FitArgs {
  const float* data;   // voxel-major [nvox][N]
  int N, nvox;
  const float* bvals;  // [N]
  const float* bvecs;  // [3][N]
  const double* pinvA; // [7][N] pseudo-inverse of the tensor design matrix
  int mode;            // FIT_PIPELINE / FIT_STAGE / FIT_START
  int model, nonlin, f0, rician;
  float Rmean;
  int kind, G, stage_f0;  // FIT_STAGE / FIT_START
  float R;
  const float* start;  // FIT_STAGE: [nvox][np]
  float* out;          // [nvox][out_np]
  int out_np;
  float* tau;          // FIT_PIPELINE with --rician
  int* iters;          // optional [nvox]
};

struct McmcArgs {
  const float* data;    // voxel-major [nvox][N]
  const float* params;  // [nvox][np] LM output
  const float* tau;     // [nvox] or null
  const float4* acq;    // [Npad] (gx, gy, gz, b), zero padded
  int N, Npad, nvox, np;
  int model, f0, rician, ardf0, no_ard, all_ard;
  float fudge, Rmean, Rstd, Rfudge;
  int nburn, njumps, se, upe, nsamples;
  uint64_t seed, vox_offset;
  // raw samples: th/ph/f [K][ns][nvox]; scalars [ns][nvox]
  float *th, *ph, *f, *S0, *d, *dstd, *R, *f0s, *taus;
  float* energy;        // [nvox] final SSR (Gaussian) / likelihood energy, optional
};

struct FinishArgs {
  int K, ns, nvox, model, f0, rician;
  const float *th, *ph, *f, *S0, *d, *dstd, *R, *f0s, *taus;  // raw samples
  float *oth, *oph, *of;  // sorted samples [K][ns][nvox]
  float *mean_S0, *mean_d, *mean_dstd, *mean_R, *mean_f0, *mean_tau, *mean_f, *dyads, *disp;
};

}  // namespace fdt
