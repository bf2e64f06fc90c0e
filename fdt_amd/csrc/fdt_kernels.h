// fdt_kernels.h -- kernel argument blocks shared by the host orchestration and
// the HIP kernels (plain data only).
#pragma once
#include <stdint.h>

namespace fdt {

enum { KIND_SINGLE = 0, KIND_SINGLE_C = 1, KIND_MULTI = 2 };
// how an LM launch obtains each voxel's start vector
enum { INIT_DTI = 0, INIT_BUF = 1, INIT_STAGE1 = 2 };

constexpr int NPMAX = 14;      // LM parameters for K <= 3
constexpr int MAX_N = 512;     // acquisition length supported
constexpr int LM_ROWS = 16;    // voxels (16-lane rows) per 256-thread LM workgroup
constexpr int NQUEUE = 8;      // work-queue counters per context

// One LM stage over a voxel work list (fdt_fit.hip).  Rows of 16 lanes take
// voxels from a device queue until it is empty.
struct LmArgs {
  const float* data;    // voxel-major [nvox][N]
  int N, nvox;
  const float4* acq;    // [N] (gx, gy, gz, b)
  const double* pinvA;  // [7][N] pseudo-inverse of the tensor design matrix
  const int* list;      // voxel indices, or null for 0..nvox-1
  const int* count;     // number of list entries (device), or null for nvox
  int* queue;           // work-queue counter, zero at launch
  int init;             // INIT_DTI (tensor start values) / INIT_BUF / INIT_STAGE1
  int lm;               // 0: write the start vector only
  int G;                // PVM_multi variant (0 gamma, 1 gamma for the ball only, 2 zeppelin)
  float R;              // zeppelin R
  int rician;           // replace non-positive samples before fitting
  const float* in;      // INIT_BUF: [vox][in_np] start vectors; INIT_STAGE1: finalised stage-1 vectors
  int in_np;
  float* out;           // [vox][out_np]: finalised parameters (or the start vector when lm == 0)
  int out_np;
  int set_f0_slot;      // also write 0.001 into the last slot (refit without f0)
  int* iters;           // optional [nvox]: LM iterations
  int iters_add;        // add to iters instead of overwriting
  const float4* acqv;   // --gradnonlin: per-voxel acquisitions [nvox][acqv_stride], or null
  int acqv_stride;
  const int* shell;     // [N] b-shell index of each volume, or null
  const float* shell_b; // [16] b value of each shell
  int nshell;           // number of shells (0: none; ignored with acqv)
};

// implausible-f0 test and refit list (CUDA/xfibres_gpu.cu:216-246, 305-339)
struct F0FlagArgs {
  const float* params;
  int nvox, np, fb;
  int K;
  int* list;
  int* count;
  unsigned char* refit;
};

// Rician precision from the residuals of the final fit (CUDA/diffmodels.cu:259-347)
struct TauArgs {
  const float* data;
  int N, nvox;
  const float4* acq;
  const float* params;
  int np, model, f0;
  float R;
  const unsigned char* refit;  // voxels refitted without f0, or null
  float* tau;
  const float4* acqv;   // --gradnonlin: per-voxel acquisitions, or null
  int acqv_stride;
};

// Chain starts as the CPU xfibres sets them without nonlinear fitting
// (cpu_start_kernel, fdt_fit.hip): the tensor start, or the means of the best
// finished neighbour when the grid is given.
struct CpuStartArgs {
  const float* data;      // voxel-major [nvox][N]
  int N, nvox;            // nvox: voxels of the volume (stride of the per-voxel arrays)
  const double* pinvA;    // [7][N] pseudo-inverse of the tensor design
  const float4* acqv;     // --gradnonlin: per-voxel acquisitions [nvox][acqv_stride], or null
  int acqv_stride;
  const int* list;        // voxels to start: list[0..count), or 0..count-1 when null
  int count;
  int model, K, np, f0;   // chain parameter layout [S0, d, (dstd), (f, th, ph) x K, (f0)]
  const int* vol2mat;     // [nz][ny][nx]: mask index + 1, 0 outside; null: tensor start only
  const long long* gidx;  // [nvox]: grid index x + nx (y + ny z) of each voxel
  int nx, ny, nz;
  const int* been;        // [nvox]: fibres with mean f > 0.05 of a finished voxel, 0 before
  const float* mean_f;    // [K][nvox], chain fibre order
  const float* mean_d;    // [nvox]
  const float* mean_S0;   // [nvox]
  const double* dyad;     // [K][3][nvox]
  float* params;          // [nvox][np] out
};

// Per-voxel means and dyads of the listed voxels' raw samples in chain fibre
// order, and their supported-fibre count (spatial_stats_kernel, fdt_finish.hip)
struct SpatialStatsArgs {
  int K, ns, vstride, count;
  const int* list;
  const float *th, *ph, *f, *S0, *d;  // raw samples: th/ph/f [K][ns][vstride], scalars [ns][vstride]
  int* been;
  float *mean_f, *mean_d, *mean_S0;
  double* dyad;
};

struct McmcArgs {
  const float* data;    // voxel-major [nvox][N]
  const float* params;  // [nvox][np] LM output
  const float* tau;     // [nvox] or null
  const float4* acq;    // [Npad] (gx, gy, gz, b), zero padded
  const int* shell;     // [Npad] b-shell of each volume (nshell > 0)
  const float* shell_b; // [16] b value of each shell
  int nshell;           // distinct b values if <= 16, else 0
  int N, Npad, nvox, np;
  int model, f0, rician, ardf0, no_ard, all_ard;
  float fudge, Rmean, Rstd, Rfudge;
  int nburn, njumps, se, upe, nsamples;
  uint64_t seed, vox_offset;
  // raw samples: th/ph/f [K][ns][nvox]; scalars [ns][nvox]
  float *th, *ph, *f, *S0, *d, *dstd, *R, *f0s, *taus;
  float* energy;        // [nvox] final SSR (Gaussian) / likelihood energy, optional
  float* ssr;           // [nvox] final sum of squared residuals y - S0 M (either noise model), optional
  const int* list;      // chains to run: voxels list[0..nvox), or 0..nvox-1 when null
  int vstride;          // voxel count of the sample arrays (>= every voxel index)
  const float* widths_in;  // [vstride][16] initial proposal width per slot (<= 0: default), optional
  float* widths_out;    // [vstride][16] final proposal widths, optional
  const float4* acqv;   // --gradnonlin: per-voxel acquisitions [nvox][acqv_stride >= Npad], or null
  int acqv_stride;
  const float* gdev;    // --gradnonlin: per-voxel deviation values [nvox][9]
};

// Batched MCMC over a queue of volumes (fdt_queue.hip): work items are runs
// of rows_per_wg consecutive voxels of one job; job j owns items
// [first_item[j], first_item[j+1]).
constexpr int QUEUE_MAX_VPL = 18;   // tiles compiled into the queue kernel: VPL 3, 5, 10, 18
struct QueueJob {
  McmcArgs a;
  int K, vpl;
};
struct QueueArgs {
  const QueueJob* jobs;
  const int* first_item;  // [njobs + 1]
  int njobs;
  int* counter;           // work-queue counter, zero at launch
};

// posterior-predictive signal from the posterior means (xfibres_pred.cc:153-190)
struct PredArgs {
  int N, nvox, K, model;
  const float* bvals;   // [N]
  const float* bvecs;   // [3][N]
  const float *S0, *d, *dstd, *R, *f0v;  // [nvox]; dstd (model >= 2), R (model 3), f0v (or null)
  const float* f;       // [K][nvox] mean f
  const float* dyads;   // [K][3][nvox]
  float* out;           // [N][nvox]
};

struct FinishArgs {
  int K, ns, nvox, model, f0, rician;
  const float *th, *ph, *f, *S0, *d, *dstd, *R, *f0s, *taus;  // raw samples
  float *oth, *oph, *of;  // sorted samples [K][ns][nvox]
  float *mean_S0, *mean_d, *mean_dstd, *mean_R, *mean_f0, *mean_tau, *mean_f, *dyads, *disp;
};

}  // namespace fdt
