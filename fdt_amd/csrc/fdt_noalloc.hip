// fdt_noalloc.hip -- the chains and the finishing pass over caller-owned
// device buffers, for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops).  Nothing here calls hipMalloc: the
// library's scratch -- the zero-padded acquisition table the chains read and
// the unsorted fibre samples the finishing pass sorts -- lives in a file-scope
// static device array, one copy per GPU, resolved once per device with
// hipGetSymbolAddress.  Every launch is stream-ordered on the caller's stream.
#include <hip/hip_runtime.h>

#include "fdt_kernels.h"
#include "fdt_xfibres.h"

#include <cstdarg>
#include <cstdio>
#include <mutex>

namespace {

using fdt::McmcArgs;
using fdt::FinishArgs;

constexpr size_t ACQ_BYTES = sizeof(float4) * 16 * 32;  // the largest chain tile: 32 volumes per lane
constexpr size_t SAMPLE_FLOATS = FDT_NOALLOC_SAMPLE_FLOATS;

// The scratch array.  Host code never passes its name to a kernel or a copy:
// on the host the name is a stand-in, not a device address; the address is
// taken per device with hipGetSymbolAddress (scratch_on).
__device__ __align__(256) unsigned char scratch[ACQ_BYTES + SAMPLE_FLOATS * sizeof(float)];

constexpr int MAX_DEVICES = 64;
std::mutex g_mtx;
unsigned char* g_base[MAX_DEVICES];  // device address of `scratch` on each device, once resolved

int fail(char* msg, size_t len, int code, const char* fmt, ...) {
  if (msg && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, len, fmt, ap);
    va_end(ap);
  }
  return code;
}

// device address of the scratch array on the current device
int scratch_on(int dev, unsigned char** out, char* msg, size_t len) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(msg, len, FDT_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_mtx);
  if (!g_base[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(msg, len, FDT_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_base[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_base[dev];
  return FDT_OK;
}

// (gx, gy, gz, b) of every slot of the chain tile, zero past N (padding
// slots must see g = 0 and b = 0)
__global__ void acq_table_kernel(const float* __restrict__ bvals, const float* __restrict__ bvecs, int N, int npad,
                                 float4* __restrict__ acq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  acq[i] = i < N ? make_float4(bvecs[i], bvecs[N + i], bvecs[2 * N + i], bvals[i]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

extern "C" {
hipError_t fdt_launch_finish(const FinishArgs*, hipStream_t);
#define FDT_NA_VPL(V) hipError_t fdt_launch_mcmc_vpl##V(const McmcArgs*, int, int, hipStream_t);
FDT_NA_VPL(3) FDT_NA_VPL(5) FDT_NA_VPL(10) FDT_NA_VPL(18) FDT_NA_VPL(24) FDT_NA_VPL(32)
#undef FDT_NA_VPL

int fdt_mcmc_noalloc(const fdt_opts* o, int N, const float* d_bvals, const float* d_bvecs, const float* d_data,
                     const float* d_params, const float* d_tau, uint64_t vox_offset, int nvox,
                     const fdt_outputs* out, void* stream, char* msg, size_t msg_len) {
  if (!o || !out) return fail(msg, msg_len, FDT_EINVAL, "null argument");
  if (o->nfibres < 1 || o->nfibres > 3) return fail(msg, msg_len, FDT_EUNSUP, "nfibres=%d: supported 1..3", o->nfibres);
  if (o->model < 1 || o->model > 3) return fail(msg, msg_len, FDT_EINVAL, "model=%d", o->model);
  if (o->nburn < 0 || o->njumps < 0 || o->sampleevery <= 0 || o->updateproposalevery <= 0)
    return fail(msg, msg_len, FDT_EINVAL, "chain lengths");
  const int K = o->nfibres, ns = o->njumps / o->sampleevery;
  if (ns <= 0) return fail(msg, msg_len, FDT_EINVAL, "njumps < sampleevery gives no samples");
  if (N < 7 || N > fdt::MAX_N) return fail(msg, msg_len, FDT_EINVAL, "N=%d: supported 7..%d", N, fdt::MAX_N);
  if (nvox < 0) return fail(msg, msg_len, FDT_EINVAL, "nvox=%d", nvox);
  if ((size_t)3 * K * ns * (size_t)nvox > SAMPLE_FLOATS)
    return fail(msg, msg_len, FDT_EINVAL, "3 x %d fibres x %d samples x %d voxels exceeds the %zu-float scratch", K, ns,
                nvox, SAMPLE_FLOATS);
  if (nvox == 0) return FDT_OK;
  const bool m2 = o->model >= 2, m3 = o->model == 3;
  if (!d_bvals || !d_bvecs || !d_data || !d_params || (o->rician && !d_tau) || !out->th || !out->ph || !out->f ||
      !out->mean_S0 || !out->mean_d || !out->mean_f || !out->dyads || !out->disp || !out->S0 || !out->d ||
      (m2 && (!out->dstd || !out->mean_dstd)) || (m3 && (!out->R || !out->mean_R)) ||
      (o->f0 && (!out->f0 || !out->mean_f0)) || (o->rician && (!out->tau || !out->mean_tau)))
    return fail(msg, msg_len, FDT_EINVAL, "a required device pointer is null");
  struct Tile { int vpl; hipError_t (*fn)(const McmcArgs*, int, int, hipStream_t); };
  const Tile tiles[] = {{3, fdt_launch_mcmc_vpl3}, {5, fdt_launch_mcmc_vpl5}, {10, fdt_launch_mcmc_vpl10},
                        {18, fdt_launch_mcmc_vpl18}, {24, fdt_launch_mcmc_vpl24}, {32, fdt_launch_mcmc_vpl32}};
  const Tile* tile = nullptr;
  for (const auto& t : tiles)
    if (16 * t.vpl >= N) { tile = &t; break; }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(msg, msg_len, FDT_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  unsigned char* base = nullptr;
  if (int rc = scratch_on(dev, &base, msg, msg_len)) return rc;
  float4* acq = reinterpret_cast<float4*>(base);
  float* raw = reinterpret_cast<float*>(base + ACQ_BYTES);
  const size_t nsv = (size_t)ns * nvox;
  hipStream_t s = static_cast<hipStream_t>(stream);

  const int npad = 16 * tile->vpl;
  hipLaunchKernelGGL(acq_table_kernel, dim3((npad + 255) / 256), dim3(256), 0, s, d_bvals, d_bvecs, N, npad, acq);
  if ((e = hipGetLastError()) != hipSuccess) return fail(msg, msg_len, FDT_EHIP, "acquisition table: %s", hipGetErrorString(e));

  McmcArgs a{};
  a.data = d_data;
  a.params = d_params;
  a.tau = o->rician ? d_tau : nullptr;
  a.acq = acq;
  a.shell = nullptr;  // no b-shell table: the isotropic signal is evaluated per volume (same values)
  a.shell_b = nullptr;
  a.nshell = 0;
  a.N = N;
  a.Npad = npad;
  a.nvox = nvox;
  a.np = 2 + 3 * K + (m2 ? 1 : 0) + (o->f0 ? 1 : 0);
  a.model = o->model;
  a.f0 = o->f0;
  a.rician = o->rician;
  a.ardf0 = o->ardf0;
  a.no_ard = o->no_ard;
  a.all_ard = o->all_ard;
  a.fudge = o->fudge;
  a.Rmean = o->R_prior_mean;
  a.Rstd = o->R_prior_std;
  a.Rfudge = o->R_prior_fudge;
  a.nburn = o->nburn;
  a.njumps = o->njumps;
  a.se = o->sampleevery;
  a.upe = o->updateproposalevery;
  a.nsamples = ns;
  a.seed = o->seed;
  a.vox_offset = vox_offset;
  a.th = raw;
  a.ph = raw + nsv * K;
  a.f = raw + 2 * nsv * K;
  a.S0 = out->S0;
  a.d = out->d;
  a.dstd = m2 ? out->dstd : nullptr;
  a.R = m3 ? out->R : nullptr;
  a.f0s = o->f0 ? out->f0 : nullptr;
  a.taus = o->rician ? out->tau : nullptr;
  a.vstride = nvox;
  if ((e = tile->fn(&a, K, o->rician, s)) != hipSuccess) return fail(msg, msg_len, FDT_EHIP, "chains: %s", hipGetErrorString(e));

  FinishArgs fa{};
  fa.K = K; fa.ns = ns; fa.nvox = nvox; fa.model = o->model; fa.f0 = o->f0; fa.rician = o->rician;
  fa.th = a.th; fa.ph = a.ph; fa.f = a.f; fa.S0 = a.S0; fa.d = a.d; fa.dstd = a.dstd; fa.R = a.R; fa.f0s = a.f0s;
  fa.taus = a.taus;
  fa.oth = out->th; fa.oph = out->ph; fa.of = out->f;
  fa.mean_S0 = out->mean_S0; fa.mean_d = out->mean_d; fa.mean_dstd = m2 ? out->mean_dstd : nullptr;
  fa.mean_R = m3 ? out->mean_R : nullptr; fa.mean_f0 = o->f0 ? out->mean_f0 : nullptr;
  fa.mean_tau = o->rician ? out->mean_tau : nullptr;
  fa.mean_f = out->mean_f; fa.dyads = out->dyads; fa.disp = out->disp;
  if ((e = fdt_launch_finish(&fa, s)) != hipSuccess) return fail(msg, msg_len, FDT_EHIP, "finish: %s", hipGetErrorString(e));
  return FDT_OK;
}

}  // extern "C"
