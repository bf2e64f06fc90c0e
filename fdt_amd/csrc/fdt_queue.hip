// fdt_queue.hip -- one persistent MCMC launch over a mixed queue of volumes.
//
// Volumes with different acquisitions (N), models and fibre counts are packed
// into one launch: a workgroup takes the next work item (a run of
// consecutive voxels of one volume, one per 16-lane row) from a device
// counter, stages that volume's acquisition in LDS and runs the chains with
// the tile of that volume (run_chain, fdt_chain.h); it takes the next item
// when all its rows are done.  Every (K, VPL) tile with VPL <= 18 is compiled
// in; the register budget is that of the largest (K=3, VPL=18).  Draws are
// keyed by each volume's global voxel index, so a voxel's chain is the same
// as in the per-volume launch with the same tile.
#include "fdt_chain.h"

namespace fdt {

template <int K, int VPL>
FDT_DEV void queue_item(const McmcArgs& a, int vox, bool active, const ChainLds& L) {
  stage_acq<VPL>(a, L);
  __syncthreads();
  if (active) run_chain<K, VPL, false>(a, vox, L);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void mcmc_queue_kernel(QueueArgs q) {
  __shared__ float4 sacq[16 * QUEUE_MAX_VPL];
  __shared__ float sba[4 * 3 * QUEUE_MAX_VPL * 64];
  __shared__ int ssh[16 * QUEUE_MAX_VPL];
  __shared__ float sshb[16];
  __shared__ float siso[16 * 16];
  __shared__ float svol[4 * QUEUE_MAX_VPL * 64];
  __shared__ int sperm[16 * QUEUE_MAX_VPL];
  __shared__ int szc[ZC_CHUNKS];
  __shared__ int sitem, sjob;
  const ChainLds L{sacq, sba, ssh, sshb, siso, nullptr, nullptr, svol, sperm, szc};
  const int rows = blockDim.x >> 4;
  const int total = q.first_item[q.njobs];
  for (;;) {
    if (threadIdx.x == 0) {
      int it = atomicAdd(q.counter, 1);
      int j = 0;
      while (j + 1 < q.njobs && q.first_item[j + 1] <= it) ++j;
      sitem = it;
      sjob = j;
    }
    __syncthreads();
    const int item = sitem, j = sjob;
    if (item >= total) break;
    const QueueJob& J = q.jobs[j];
    const int vox = (item - q.first_item[j]) * rows + (threadIdx.x >> 4);
    const bool active = vox < J.a.nvox;
    const int variant = (J.K - 1) * 4 + (J.vpl == 3 ? 0 : J.vpl == 5 ? 1 : J.vpl == 10 ? 2 : 3);
    switch (variant) {
      case 0: queue_item<1, 3>(J.a, vox, active, L); break;
      case 1: queue_item<1, 5>(J.a, vox, active, L); break;
      case 2: queue_item<1, 10>(J.a, vox, active, L); break;
      case 3: queue_item<1, 18>(J.a, vox, active, L); break;
      case 4: queue_item<2, 3>(J.a, vox, active, L); break;
      case 5: queue_item<2, 5>(J.a, vox, active, L); break;
      case 6: queue_item<2, 10>(J.a, vox, active, L); break;
      case 7: queue_item<2, 18>(J.a, vox, active, L); break;
      case 8: queue_item<3, 3>(J.a, vox, active, L); break;
      case 9: queue_item<3, 5>(J.a, vox, active, L); break;
      case 10: queue_item<3, 10>(J.a, vox, active, L); break;
      default: queue_item<3, 18>(J.a, vox, active, L); break;
    }
    __syncthreads();
  }
}

}  // namespace fdt

// rows_per_wg voxels per workgroup (4, 8 or 16: 64, 128 or 256 threads);
// grid: as many workgroups as stay resident
extern "C" hipError_t fdt_launch_mcmc_queue(const fdt::QueueArgs* q, int rows_per_wg, int max_items, hipStream_t s) {
  if (max_items <= 0) return hipSuccess;
  if (rows_per_wg != 4 && rows_per_wg != 8 && rows_per_wg != 16) return hipErrorInvalidValue;
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fdt::mcmc_queue_kernel, 16 * rows_per_wg, 0)) != hipSuccess)
    return e;
  long cap = (long)(per_cu > 0 ? per_cu : 1) * cus;
  int grid = (int)(max_items < cap ? max_items : cap);
  hipLaunchKernelGGL(fdt::mcmc_queue_kernel, dim3(grid), dim3(16 * rows_per_wg), 0, s, *q);
  return hipGetLastError();
}
