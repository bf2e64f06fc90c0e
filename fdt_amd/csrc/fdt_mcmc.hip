// fdt_mcmc.hip -- the persistent MCMC launch over one volume: one chain per
// 16-lane row, 16 chains per 256-thread workgroup, the whole nburn + njumps
// loop in one launch.  The chain itself is run_chain (fdt_chain.h).
#include "fdt_chain.h"

namespace fdt {

// XCD-aware tile order.  The dispatcher deals workgroups round-robin over the
// 8 XCDs, each with its own L2; renumbered this way each XCD runs one
// contiguous range of tiles, so the 64-byte record runs of neighbouring
// workgroups (16 voxels x 4 B per sample and quantity) meet in one L2 and
// leave it as whole lines.  A bijection of [0, nb) for every nb.
__device__ inline int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, l = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}

template <int K, int VPL, bool RIC, bool GN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VPL <= 18 ? 2 : 1, 2))) void mcmc_kernel(McmcArgs a) {
  __shared__ float4 sacq[16 * VPL];
  __shared__ float sba[4 * K * VPL * 64];
  __shared__ int ssh[GN ? 1 : 16 * VPL];
  __shared__ float sshb[16];
  __shared__ float siso[GN ? 1 : 16 * 16];
  __shared__ float sbp[GN ? 16 * 16 * VPL : 1];
  __shared__ float sgd[GN ? 16 * 9 : 1];
  __shared__ float svol[GN ? 1 : 4 * VPL * 64];
  constexpr bool ZS = ZERO_SLOT<VPL>;
  __shared__ int sperm[ZS ? 16 * VPL : 1];
  __shared__ int szc[ZS ? ZC_CHUNKS : 1];
  const ChainLds L{sacq, sba, GN ? nullptr : ssh, sshb, siso, sbp, sgd, GN ? nullptr : svol, ZS ? sperm : nullptr,
                   ZS ? szc : nullptr};
  stage_acq<VPL>(a, L);
  __syncthreads();
  const int i = xcd_tile(blockIdx.x, gridDim.x) * 16 + (threadIdx.x >> 4);
  if (i < a.nvox) run_chain<K, VPL, RIC, GN>(a, a.list ? a.list[i] : i, L);
}

}  // namespace fdt

// Two translation units per VPL value (-DFDT_VPL=...): -DFDT_MCMC_RIC=0 holds
// the Gaussian kernels and the launcher, -DFDT_MCMC_RIC=1 the Rician ones
// (built in parallel; each can take its own compiler options).  K = 1..3, with the shared
// acquisition or per-voxel ones (--gradnonlin, a->acqv set; its stride must
// cover 16 * VPL volumes).
#ifndef FDT_VPL
#define FDT_VPL 18
#endif
#ifndef FDT_MCMC_RIC
#define FDT_MCMC_RIC 0
#endif
#define FDT_CAT2_(a, b) a##b
#define FDT_CAT2(a, b) FDT_CAT2_(a, b)
#define FDT_MC_CASE(KK, RR, GG) \
  if (K == KK && gn == GG) { hipLaunchKernelGGL((fdt::mcmc_kernel<KK, FDT_VPL, RR, GG>), grid, block, 0, s, *a); return hipGetLastError(); }

#if FDT_MCMC_RIC
extern "C" hipError_t FDT_CAT2(fdt_launch_mcmc_ric_vpl, FDT_VPL)(const fdt::McmcArgs* a, int K, hipStream_t s) {
  dim3 grid((a->nvox + 15) / 16), block(256);
  const bool gn = a->acqv != nullptr;
  FDT_MC_CASE(1, true, false) FDT_MC_CASE(2, true, false) FDT_MC_CASE(3, true, false)
  FDT_MC_CASE(1, true, true) FDT_MC_CASE(2, true, true) FDT_MC_CASE(3, true, true)
  return hipErrorInvalidValue;
}
#else
extern "C" hipError_t FDT_CAT2(fdt_launch_mcmc_ric_vpl, FDT_VPL)(const fdt::McmcArgs* a, int K, hipStream_t s);
extern "C" hipError_t FDT_CAT2(fdt_launch_mcmc_vpl, FDT_VPL)(const fdt::McmcArgs* a, int K, int rician, hipStream_t s) {
  if (a->nvox <= 0) return hipSuccess;
  if (a->acqv && a->acqv_stride < 16 * FDT_VPL) return hipErrorInvalidValue;
  if (rician) return FDT_CAT2(fdt_launch_mcmc_ric_vpl, FDT_VPL)(a, K, s);
  dim3 grid((a->nvox + 15) / 16), block(256);
  const bool gn = a->acqv != nullptr;
  FDT_MC_CASE(1, false, false) FDT_MC_CASE(2, false, false) FDT_MC_CASE(3, false, false)
  FDT_MC_CASE(1, false, true) FDT_MC_CASE(2, false, true) FDT_MC_CASE(3, false, true)
  return hipErrorInvalidValue;
}
#endif
#undef FDT_MC_CASE
