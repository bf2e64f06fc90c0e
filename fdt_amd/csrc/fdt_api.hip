// fdt_api.hip -- host orchestration behind the C ABI of include/fdt_xfibres.h.
//
// Replaces the reference's per-sub-part host pipeline (xfibres_gpu(),
// CUDA/xfibres_gpu.cu:29-146) with a context that owns one stream and the
// device buffers; data, LM parameters and samples stay resident between
// calls.  No call exits the process (the reference's sync_check does,
// CUDA/sync_check.h:18-25); errors come back as FDT_E* codes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <random>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fdt_xfibres.h"
#include "fdt_kernels.h"
#include "fdt_mlp.h"

using namespace fdt;

extern "C" {
#define FDT_LM_DECL(K, P)                                                              \
  hipError_t fdt_launch_lm_k##K##_##P(const LmArgs*, int kind, int f0, hipStream_t); \
  hipError_t fdt_launch_tau_k##K##_##P(const TauArgs*, hipStream_t);
FDT_LM_DECL(1, 0)
FDT_LM_DECL(2, 0)
FDT_LM_DECL(3, 0)
FDT_LM_DECL(1, 1)
FDT_LM_DECL(2, 1)
FDT_LM_DECL(3, 1)
FDT_LM_DECL(1, 2)
FDT_LM_DECL(2, 2)
FDT_LM_DECL(3, 2)
hipError_t fdt_launch_f0_flag(const F0FlagArgs*, hipStream_t);
hipError_t fdt_launch_zero_ints(int*, int, hipStream_t);
hipError_t fdt_launch_mcmc_queue(const QueueArgs*, int rows_per_wg, int max_items, hipStream_t);
hipError_t fdt_launch_gemm_bf16(const GemmArgs*, hipStream_t);
hipError_t fdt_launch_mlp_ew(int op, const void* a, const void* b, void* c, size_t n, int M, int N, hipStream_t);
hipError_t fdt_launch_adam(float* w, float* m, float* v, const float* g, uint16_t* wb, size_t n, float lr, float b1,
                           float b2, float eps, int step, hipStream_t);
hipError_t fdt_launch_normalize(const float* data, int N, int nvox, const float* bvals, float b0_max, uint16_t* X,
                                float* S0, hipStream_t);
hipError_t fdt_launch_sg_loss(const float* O, const float* T, int M, uint16_t* dO, float* loss_part, float w_lw,
                              hipStream_t);
hipError_t fdt_launch_sg_decode(const SgDecodeArgs*, hipStream_t);
hipError_t fdt_launch_sg_fallback(const SgFallbackArgs*, hipStream_t);
hipError_t fdt_launch_finish(const FinishArgs*, hipStream_t);
hipError_t fdt_launch_cpu_start(const CpuStartArgs*, hipStream_t);
hipError_t fdt_launch_spatial_stats(const SpatialStatsArgs*, hipStream_t);
hipError_t fdt_launch_transpose(const float*, float*, int, int, hipStream_t);
hipError_t fdt_launch_pred(const PredArgs*, hipStream_t);
hipError_t fdt_launch_grad_acq(const float4* base, int N, int stride, const float* grad, int nvox, float4* out,
                               hipStream_t s);
#define FDT_VPL_DECL(V) hipError_t fdt_launch_mcmc_vpl##V(const McmcArgs*, int, int, hipStream_t);
FDT_VPL_DECL(3)
FDT_VPL_DECL(5)
FDT_VPL_DECL(10)
FDT_VPL_DECL(18)
FDT_VPL_DECL(24)
FDT_VPL_DECL(32)
}

namespace {

typedef hipError_t (*mcmc_launch_fn)(const McmcArgs*, int, int, hipStream_t);
struct VplEntry { int vpl; mcmc_launch_fn fn; };
const VplEntry kVpl[] = {{3, fdt_launch_mcmc_vpl3}, {5, fdt_launch_mcmc_vpl5}, {10, fdt_launch_mcmc_vpl10},
                         {18, fdt_launch_mcmc_vpl18}, {24, fdt_launch_mcmc_vpl24}, {32, fdt_launch_mcmc_vpl32}};

// the MCMC tile (volumes per lane) for N volumes, at least `force`
int tile_of(int N, int force) {
  for (const auto& e : kVpl)
    if (16 * e.vpl >= N && e.vpl >= force) return e.vpl;
  return 0;
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t want) {
    if (want <= n && p) return hipSuccess;
    if (p) { (void)hipFree(p); p = nullptr; n = 0; }
    if (want == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, want * sizeof(T));
    if (e == hipSuccess) n = want;
    return e;
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// Pseudo-inverse (7 x N) of the tensor design matrix (DTI::form_Amat,
// diffmodels.h:148-163) by eigen-decomposition of the column-equilibrated
// normal matrix; singular values below max(N,7)*s_max*1e-15 are dropped.
void jacobi_sym(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& V) {
  V.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) { tot += A[i * n + j] * A[i * n + j]; if (i != j) off += A[i * n + j] * A[i * n + j]; }
    if (off == 0 || off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (apq == 0) continue;
        double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) { double a = A[k * n + p], b = A[k * n + q]; A[k * n + p] = c * a - s * b; A[k * n + q] = s * a + c * b; }
        for (int k = 0; k < n; ++k) { double a = A[p * n + k], b = A[q * n + k]; A[p * n + k] = c * a - s * b; A[q * n + k] = s * a + c * b; }
        for (int k = 0; k < n; ++k) { double a = V[k * n + p], b = V[k * n + q]; V[k * n + p] = c * a - s * b; V[k * n + q] = s * a + c * b; }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

std::vector<double> dti_pinv(int N, const std::vector<float>& bvals, const std::vector<float>& bvecs) {
  const int n = 7;
  std::vector<double> A((size_t)N * n);
  for (int i = 0; i < N; ++i) {
    double x = bvecs[i], y = bvecs[N + i], z = bvecs[2 * N + i], b = bvals[i];
    double* r = &A[(size_t)i * n];
    r[0] = x * x * b; r[1] = 2 * x * y * b; r[2] = 2 * x * z * b;
    r[3] = y * y * b; r[4] = 2 * y * z * b; r[5] = z * z * b; r[6] = 1.0;
  }
  // equilibrate columns: A = As * D, pinv(A) = D^-1 pinv(As) for full rank
  std::vector<double> scale(n, 1.0);
  for (int j = 0; j < n; ++j) {
    double s = 0;
    for (int i = 0; i < N; ++i) s += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    scale[j] = s > 0 ? std::sqrt(s) : 1.0;
  }
  std::vector<double> G(n * n, 0.0), w, V;
  for (int a = 0; a < n; ++a)
    for (int c = 0; c < n; ++c) {
      double s = 0;
      for (int i = 0; i < N; ++i) s += (A[(size_t)i * n + a] / scale[a]) * (A[(size_t)i * n + c] / scale[c]);
      G[a * n + c] = s;
    }
  jacobi_sym(n, G, w, V);
  double smax = 0;
  for (double x : w) smax = std::max(smax, std::sqrt(std::max(x, 0.0)));
  double tol = std::max(N, n) * smax * 1e-15;
  std::vector<double> inv(n);
  for (int j = 0; j < n; ++j) { double s = std::sqrt(std::max(w[j], 0.0)); inv[j] = s > tol ? 1.0 / (s * s) : 0.0; }
  std::vector<double> Gi(n * n, 0.0);
  for (int a = 0; a < n; ++a)
    for (int c = 0; c < n; ++c) {
      double s = 0;
      for (int k = 0; k < n; ++k) s += V[a * n + k] * inv[k] * V[c * n + k];
      Gi[a * n + c] = s;
    }
  std::vector<double> P((size_t)n * N);
  for (int a = 0; a < n; ++a)
    for (int i = 0; i < N; ++i) {
      double s = 0;
      for (int c = 0; c < n; ++c) s += Gi[a * n + c] * (A[(size_t)i * n + c] / scale[c]);
      P[(size_t)a * N + i] = s / scale[a];
    }
  return P;
}

}  // namespace

struct fdt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  fdt_opts o{};
  std::string err;
  int N = 0;
  std::vector<float> bvals, bvecs;
  DevBuf<float> d_bvals, d_bvecs;
  DevBuf<double> d_pinv;
  DevBuf<float4> d_acq;
  DevBuf<int> d_shell;
  DevBuf<float> d_shell_b;
  int nshell = 0;
  int Npad = 0;
  int nvox = 0;
  DevBuf<float> d_data, d_stage, d_params, d_tau, d_energy;
  DevBuf<int> d_iters, d_queue, d_list;
  DevBuf<float> d_s1;             // finalised stage-1 vectors of the model >= 2 pipeline
  DevBuf<float> d_widths_in, d_widths_out;  // proposal widths per voxel and slot [nvox][16]
  bool use_widths_in = false;
  bool energy_valid = false;  // d_energy holds the final energy of the last chains
  DevBuf<unsigned char> d_refit;  // voxels refitted without f0
  bool have_params = false, have_samples = false, have_finish = false;
  bool have_iters = false;  // d_iters holds the LM iterations of the current parameters
  // raw and sorted samples
  DevBuf<float> s_th, s_ph, s_f, s_S0, s_d, s_dstd, s_R, s_f0, s_tau;
  DevBuf<float> o_th, o_ph, o_f;
  DevBuf<float> m_S0, m_d, m_dstd, m_R, m_f0, m_tau, m_f, m_dyads, m_disp;
  hipEvent_t ev[6] = {};
  int vpl_force = 0, vpl_used = 0;
  bool timed_fit = false, timed_mcmc = false, timed_finish = false;
  bool quiet = false;  // no timing events (launches recorded into a graph)
  uint64_t vox_offset = 0;
  hipEvent_t ev_done = nullptr;
  // --gradnonlin: per-voxel gradient deviations [grad_nvox][9] and the corrected
  // acquisitions built from them [nvox][acqv_stride] (rebuilt when stale)
  bool gn = false, acqv_valid = false;
  int grad_nvox = 0, acqv_stride = 0;
  DevBuf<float> d_grad;
  DevBuf<float4> d_acqv;
  // the voxel grid (fdt_set_grid) and the per-voxel state of the CPU xfibres
  // neighbour initialisation, in chain fibre order
  int gnx = 0, gny = 0, gnz = 0;
  std::vector<long long> h_gidx;
  std::vector<int> h_wlist;  // wavefront lists of the last neighbour-initialised run
  DevBuf<long long> d_gidx;
  DevBuf<int> d_vol2mat, d_been, d_wlist;
  DevBuf<float> sp_f, sp_d, sp_S0;
  DevBuf<double> sp_dyad;
  bool have_spatial = false;
};

static int set_err(fdt_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}
#define HIPCHK(c, expr)                                                                           \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return set_err(c, FDT_EHIP, "%s: %s", #expr, hipGetErrorString(e_));    \
  } while (0)

static int nsamples_of(const fdt_opts& o) { return o.sampleevery > 0 ? o.njumps / o.sampleevery : 0; }
static int np_of(const fdt_opts& o) { return 2 + 3 * o.nfibres + (o.model >= 2 ? 1 : 0) + (o.f0 ? 1 : 0); }

static int check_opts(fdt_ctx* c, const fdt_opts& o) {
  if (o.nfibres < 1 || o.nfibres > 3) return set_err(c, FDT_EUNSUP, "nfibres=%d: this build supports 1..3", o.nfibres);
  if (o.model < 1 || o.model > 3) return set_err(c, FDT_EINVAL, "model=%d: must be 1, 2 or 3", o.model);
  if (o.njumps < 0 || o.nburn < 0) return set_err(c, FDT_EINVAL, "negative jump counts");
  if (o.sampleevery <= 0) return set_err(c, FDT_EINVAL, "sampleevery must be > 0");
  if (o.updateproposalevery <= 0) return set_err(c, FDT_EINVAL, "updateproposalevery must be > 0");
  if (o.njumps > 0 && nsamples_of(o) == 0) return set_err(c, FDT_EINVAL, "njumps < sampleevery gives no samples");
  if (o.lm_fp64 < 0 || o.lm_fp64 > 2) return set_err(c, FDT_EINVAL, "lm_fp64=%d: must be 0, 1 or 2", o.lm_fp64);
  return FDT_OK;
}

static int ensure_acqv(fdt_ctx* c);

extern "C" {

fdt_ctx* fdt_create(int device, const fdt_opts* o) {
  if (!o) return nullptr;
  fdt_ctx* c = new fdt_ctx;
  c->device = device;
  c->o = *o;
  if (check_opts(c, *o) != FDT_OK) { delete c; return nullptr; }
  if (hipSetDevice(device) != hipSuccess) { delete c; return nullptr; }
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return nullptr; }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) { delete c; return nullptr; }
  return c;
}

void fdt_destroy(fdt_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  DevBuf<float>* fb[] = {&c->d_bvals, &c->d_bvecs, &c->d_data, &c->d_stage, &c->d_params, &c->d_tau, &c->d_energy,
                         &c->s_th, &c->s_ph, &c->s_f, &c->s_S0, &c->s_d, &c->s_dstd, &c->s_R, &c->s_f0, &c->s_tau,
                         &c->o_th, &c->o_ph, &c->o_f, &c->m_S0, &c->m_d, &c->m_dstd, &c->m_R, &c->m_f0, &c->m_tau,
                         &c->m_f, &c->m_dyads, &c->m_disp};
  for (auto* b : fb) b->release();
  c->d_pinv.release();
  c->d_acq.release();
  c->d_grad.release();
  c->d_acqv.release();
  c->d_shell.release();
  c->d_shell_b.release();
  c->d_iters.release();
  c->d_queue.release();
  c->d_list.release();
  c->d_s1.release();
  c->d_refit.release();
  c->d_widths_in.release();
  c->d_widths_out.release();
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  if (c->ev_done) (void)hipEventDestroy(c->ev_done);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* fdt_last_error(const fdt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fdt_num_params(const fdt_ctx* c) { return c ? np_of(c->o) : FDT_EINVAL; }

int fdt_set_acquisition(fdt_ctx* c, int N, const float* bvals, const float* bvecs) {
  if (!c || !bvals || !bvecs) return set_err(c, FDT_EINVAL, "null argument");
  if (N < 7 || N > MAX_N) return set_err(c, FDT_EINVAL, "N=%d: supported range is 7..%d volumes", N, MAX_N);
  (void)hipSetDevice(c->device);
  c->N = N;
  c->bvals.assign(bvals, bvals + N);
  c->bvecs.assign(bvecs, bvecs + 3 * N);
  std::vector<double> P = dti_pinv(N, c->bvals, c->bvecs);
  int vpl = 0;
  for (const auto& e : kVpl)
    if (16 * e.vpl >= N) { vpl = e.vpl; break; }
  if (!vpl) return set_err(c, FDT_EUNSUP, "N=%d exceeds the largest MCMC tile", N);
  if (c->vpl_force) vpl = std::max(vpl, c->vpl_force);
  c->Npad = 16 * 32;  // room for every tile
  std::vector<float4> acq(c->Npad, make_float4(0, 0, 0, 0));
  for (int i = 0; i < N; ++i) acq[i] = make_float4(bvecs[i], bvecs[N + i], bvecs[2 * N + i], bvals[i]);
  HIPCHK(c, c->d_bvals.ensure(N));
  HIPCHK(c, c->d_bvecs.ensure(3 * N));
  HIPCHK(c, c->d_pinv.ensure((size_t)7 * N));
  HIPCHK(c, c->d_acq.ensure(c->Npad));
  HIPCHK(c, hipMemcpyAsync(c->d_bvals.p, bvals, sizeof(float) * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_bvecs.p, bvecs, sizeof(float) * 3 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_pinv.p, P.data(), sizeof(double) * 7 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_acq.p, acq.data(), sizeof(float4) * c->Npad, hipMemcpyHostToDevice, c->stream));
  // b-shells: distinct b values (exact), used by the MCMC when there are at most 16.
  // Padding slots (i >= N) belong to a b = 0 shell, as their acquisition is
  // zero: the chain relies on their signals being exactly 1.
  std::vector<float> shb;
  std::vector<int> shell(c->Npad, 0);
  for (int i = 0; i < c->Npad && shb.size() <= 16; ++i) {
    const float b = i < N ? bvals[i] : 0.0f;
    auto it = std::find(shb.begin(), shb.end(), b);
    if (it == shb.end()) { shb.push_back(b); shell[i] = (int)shb.size() - 1; }
    else shell[i] = (int)(it - shb.begin());
  }
  c->nshell = shb.size() <= 16 ? (int)shb.size() : 0;
  shb.resize(16, 0.0f);
  HIPCHK(c, c->d_shell.ensure(c->Npad));
  HIPCHK(c, c->d_shell_b.ensure(16));
  HIPCHK(c, hipMemcpyAsync(c->d_shell.p, shell.data(), sizeof(int) * c->Npad, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_shell_b.p, shb.data(), sizeof(float) * 16, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nvox = 0;
  c->acqv_valid = false;
  c->have_params = c->have_samples = c->have_finish = false;
  return FDT_OK;
}

int fdt_load_data(fdt_ctx* c, const float* data, int layout, int nvox) {
  if (!c) return FDT_EINVAL;
  if (c->N <= 0) return set_err(c, FDT_ESTATE, "fdt_set_acquisition must come first");
  if (nvox < 0 || (nvox > 0 && !data)) return set_err(c, FDT_EINVAL, "bad data/nvox");
  if (layout != FDT_LAYOUT_VOX_MAJOR && layout != FDT_LAYOUT_VOL_MAJOR) return set_err(c, FDT_EINVAL, "bad layout %d", layout);
  (void)hipSetDevice(c->device);
  size_t n = (size_t)nvox * c->N;
  HIPCHK(c, c->d_data.ensure(std::max<size_t>(n, 1)));
  if (n) {
    if (layout == FDT_LAYOUT_VOX_MAJOR) {
      HIPCHK(c, hipMemcpyAsync(c->d_data.p, data, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    } else {
      HIPCHK(c, c->d_stage.ensure(n));
      HIPCHK(c, hipMemcpyAsync(c->d_stage.p, data, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, fdt_launch_transpose(c->d_stage.p, c->d_data.p, c->N, nvox, c->stream));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->d_stage.release();
  c->nvox = nvox;
  c->acqv_valid = false;
  c->have_params = c->have_samples = c->have_finish = c->have_spatial = false;
  c->use_widths_in = false;
  return FDT_OK;
}

int fdt_load_grad_dev(fdt_ctx* c, const float* grad, int layout, int nvox) {
  if (!c) return FDT_EINVAL;
  if (!grad) {  // back to the shared acquisition
    c->gn = c->acqv_valid = false;
    c->grad_nvox = 0;
    c->d_grad.release();
    c->d_acqv.release();
    c->have_params = c->have_samples = c->have_finish = false;
    return FDT_OK;
  }
  if (nvox <= 0) return set_err(c, FDT_EINVAL, "nvox=%d", nvox);
  if (layout != FDT_LAYOUT_VOX_MAJOR && layout != FDT_LAYOUT_VOL_MAJOR) return set_err(c, FDT_EINVAL, "bad layout %d", layout);
  (void)hipSetDevice(c->device);
  size_t n = (size_t)nvox * 9;
  HIPCHK(c, c->d_grad.ensure(n));
  if (layout == FDT_LAYOUT_VOX_MAJOR) {
    HIPCHK(c, hipMemcpyAsync(c->d_grad.p, grad, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHK(c, c->d_stage.ensure(n));
    HIPCHK(c, hipMemcpyAsync(c->d_stage.p, grad, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, fdt_launch_transpose(c->d_stage.p, c->d_grad.p, 9, nvox, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->d_stage.release();
  c->gn = true;
  c->acqv_valid = false;
  c->grad_nvox = nvox;
  c->have_params = c->have_samples = c->have_finish = false;
  return FDT_OK;
}

int fdt_get_voxel_acquisition(fdt_ctx* c, int vox, float* bvals, float* bvecs) {
  if (!c || !bvals || !bvecs) return set_err(c, FDT_EINVAL, "null argument");
  if (vox < 0 || vox >= c->nvox) return set_err(c, FDT_EINVAL, "voxel %d of %d", vox, c->nvox);
  (void)hipSetDevice(c->device);
  const int N = c->N;
  if (!c->gn) {
    std::copy(c->bvals.begin(), c->bvals.end(), bvals);
    std::copy(c->bvecs.begin(), c->bvecs.end(), bvecs);
    return FDT_OK;
  }
  if (int rc = ensure_acqv(c)) return rc;
  std::vector<float4> row(N);
  HIPCHK(c, hipMemcpyAsync(row.data(), c->d_acqv.p + (size_t)vox * c->acqv_stride, N * sizeof(float4),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < N; ++i) {
    bvals[i] = row[i].w;
    bvecs[i] = row[i].x; bvecs[N + i] = row[i].y; bvecs[2 * N + i] = row[i].z;
  }
  return FDT_OK;
}

// prec: 0 FP32, 1 FP64, 2 FP32 normal equations with FP64 parameters and cost
static hipError_t launch_lm(const LmArgs& a, int K, int prec, int kind, int f0, hipStream_t s) {
  if (prec == 0) {
    if (K == 1) return fdt_launch_lm_k1_0(&a, kind, f0, s);
    if (K == 2) return fdt_launch_lm_k2_0(&a, kind, f0, s);
    if (K == 3) return fdt_launch_lm_k3_0(&a, kind, f0, s);
  } else if (prec == 1) {
    if (K == 1) return fdt_launch_lm_k1_1(&a, kind, f0, s);
    if (K == 2) return fdt_launch_lm_k2_1(&a, kind, f0, s);
    if (K == 3) return fdt_launch_lm_k3_1(&a, kind, f0, s);
  } else if (prec == 2) {
    if (K == 1) return fdt_launch_lm_k1_2(&a, kind, f0, s);
    if (K == 2) return fdt_launch_lm_k2_2(&a, kind, f0, s);
    if (K == 3) return fdt_launch_lm_k3_2(&a, kind, f0, s);
  }
  return hipErrorInvalidValue;
}

static hipError_t launch_tau(const TauArgs& a, int K, int prec, hipStream_t s) {
  if (prec == 0) {
    if (K == 1) return fdt_launch_tau_k1_0(&a, s);
    if (K == 2) return fdt_launch_tau_k2_0(&a, s);
    if (K == 3) return fdt_launch_tau_k3_0(&a, s);
  } else if (prec == 1) {
    if (K == 1) return fdt_launch_tau_k1_1(&a, s);
    if (K == 2) return fdt_launch_tau_k2_1(&a, s);
    if (K == 3) return fdt_launch_tau_k3_1(&a, s);
  } else if (prec == 2) {
    if (K == 1) return fdt_launch_tau_k1_2(&a, s);
    if (K == 2) return fdt_launch_tau_k2_2(&a, s);
    if (K == 3) return fdt_launch_tau_k3_2(&a, s);
  }
  return hipErrorInvalidValue;
}

// --gradnonlin: (re)build the per-voxel acquisition table for the current
// acquisition, deviations and MCMC tile (stride = 16 x the tile's volumes per lane)
static int ensure_acqv(fdt_ctx* c) {
  if (!c->gn) return FDT_OK;
  if (c->grad_nvox != c->nvox)
    return set_err(c, FDT_ESTATE, "gradient deviations for %d voxels, data for %d", c->grad_nvox, c->nvox);
  int stride = 16 * tile_of(c->N, c->vpl_force);
  if (stride <= 0) return set_err(c, FDT_EUNSUP, "no MCMC tile for N=%d", c->N);
  if (c->acqv_valid && stride == c->acqv_stride) return FDT_OK;
  HIPCHK(c, c->d_acqv.ensure((size_t)std::max(c->nvox, 1) * stride));
  HIPCHK(c, fdt_launch_grad_acq(c->d_acq.p, c->N, stride, c->d_grad.p, c->nvox, c->d_acqv.p, c->stream));
  c->acqv_stride = stride;
  c->acqv_valid = true;
  return FDT_OK;
}

static LmArgs base_lm_args(fdt_ctx* c) {
  LmArgs a{};
  a.data = c->d_data.p;
  a.N = c->N;
  a.nvox = c->nvox;
  a.acq = c->d_acq.p;
  a.pinvA = c->d_pinv.p;
  a.lm = 1;
  a.R = c->o.R_prior_mean;
  a.acqv = c->gn ? c->d_acqv.p : nullptr;
  a.acqv_stride = c->acqv_stride;
  a.shell = c->d_shell.p;
  a.shell_b = c->d_shell_b.p;
  a.nshell = c->gn ? 0 : c->nshell;
  return a;
}

// The CPU xfibres initialisation in force (xfibresVoxelManager::initialise,
// xfibres.cc:517-541): 0 the xfibres_gpu fit (also for --rician, which the CPU
// path always fits nonlinearly, and for --nonlinear / --cnonlinear), 1 the
// tensor start (--nospat), 2 the best finished neighbour, else the tensor.
static int cpu_mode(const fdt_opts& o) {
  if (!o.cpu_init || o.rician || o.nonlin || o.cnonlin) return 0;
  return o.nospat ? 1 : 2;
}

static CpuStartArgs cpu_start_args(fdt_ctx* c) {
  CpuStartArgs a{};
  a.data = c->d_data.p;
  a.N = c->N;
  a.nvox = c->nvox;
  a.pinvA = c->d_pinv.p;
  a.acqv = c->gn ? c->d_acqv.p : nullptr;
  a.acqv_stride = c->acqv_stride;
  a.model = c->o.model;
  a.K = c->o.nfibres;
  a.np = np_of(c->o);
  a.f0 = c->o.f0;
  a.params = c->d_params.p;
  return a;
}

static int cpu_mode_check(fdt_ctx* c, int mode) {
  if (c->o.model == 3)
    return set_err(c, FDT_EUNSUP, "model 3 without nonlinear fitting: initialise_tensor sets no d for model 3 "
                                  "(xfibres.cc:548-557), which leaves the chain at d = 0");
  if (mode == 2 && c->o.model != 1)
    return set_err(c, FDT_EUNSUP, "neighbour initialisation for model %d: neighbour_initialise sets no d_std "
                                  "(xfibres.cc:338-340), which leaves it at 0; use nospat for the tensor start",
                   c->o.model);
  return FDT_OK;
}

// tensor start of every voxel (initialise_tensor, xfibres.cc:546-572)
static int tensor_start_async(fdt_ctx* c) {
  if (int rc = cpu_mode_check(c, 1)) return rc;
  size_t nv = std::max(c->nvox, 1);
  HIPCHK(c, c->d_params.ensure(nv * np_of(c->o)));
  HIPCHK(c, c->d_tau.ensure(nv));
  if (int rc = ensure_acqv(c)) return rc;
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  HIPCHK(c, fdt_launch_zero_ints(reinterpret_cast<int*>(c->d_tau.p), c->nvox, c->stream));  // no tau: Gaussian noise
  CpuStartArgs a = cpu_start_args(c);
  a.count = c->nvox;
  HIPCHK(c, fdt_launch_cpu_start(&a, c->stream));
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  c->timed_fit = !c->quiet;
  c->have_params = true; c->have_iters = false;
  c->have_samples = c->have_finish = false;
  return FDT_OK;
}

// The fit() pipeline of CUDA/xfibres_gpu.cu:173-348 as a chain of launches
// on the context stream: stage 1 (PVM_single, or PVM_single_c for model 1
// without --nonlinear), PVM_multi for model >= 2, the --f0 refit of
// implausible voxels over a device list (model 1: the same stage without f0;
// model >= 2: PVM_single_c then PVM_multi, :331-333), and tau for --rician.
static int fit_async(fdt_ctx* c) {
  if (c->N <= 0) return set_err(c, FDT_ESTATE, "no acquisition");
  const fdt_opts& o = c->o;
  if (const int mode = cpu_mode(o)) {
    if (mode == 2)
      return set_err(c, FDT_ESTATE, "neighbour initialisation runs inside fdt_run (each start needs the chains of "
                                    "the neighbours before it); set nospat for the tensor start");
    return tensor_start_async(c);
  }
  // --f0 fits run the FP64 normal equations in the mixed mode: the f0 column
  // is nearly collinear with the isotropic ones and the FP32 Hessian steers
  // them off the FP64 fits (within 1e-4 of the FP64 oracle on 45% of voxels
  // against 87% on the multi-shell volume)
  const int K = o.nfibres, f0 = o.f0 ? 1 : 0, fp64 = (o.lm_fp64 == 2 && f0) ? 1 : o.lm_fp64;
  const int np = np_of(o), np1 = 2 + 3 * K + f0;
  size_t nv = std::max(c->nvox, 1);
  HIPCHK(c, c->d_params.ensure(nv * np));
  HIPCHK(c, c->d_tau.ensure(nv));
  HIPCHK(c, c->d_iters.ensure(nv));
  HIPCHK(c, c->d_queue.ensure(NQUEUE));
  if (o.model >= 2) HIPCHK(c, c->d_s1.ensure(nv * np1));
  if (f0) {
    HIPCHK(c, c->d_list.ensure(nv));
    HIPCHK(c, c->d_refit.ensure(nv));
  }
  int* q = c->d_queue.p;
  if (int rc = ensure_acqv(c)) return rc;
  const int k1 = o.model == 1 ? (o.nonlin ? KIND_SINGLE : KIND_SINGLE_C) : KIND_SINGLE;
  const int G = o.model == 3 ? 2 : 0;
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  HIPCHK(c, fdt_launch_zero_ints(q, NQUEUE, c->stream));
  LmArgs a = base_lm_args(c);
  a.rician = o.rician;
  a.iters = c->d_iters.p;
  // stage 1 from the tensor start values
  a.init = INIT_DTI;
  a.queue = q + 0;
  a.out = o.model == 1 ? c->d_params.p : c->d_s1.p;
  a.out_np = o.model == 1 ? np : np1;
  a.iters_add = 0;
  HIPCHK(c, launch_lm(a, K, fp64, k1, f0, c->stream));
  if (o.model >= 2) {
    LmArgs b = a;
    b.init = INIT_STAGE1;
    b.in = c->d_s1.p;
    b.in_np = np1;
    b.G = G;
    b.queue = q + 1;
    b.out = c->d_params.p;
    b.out_np = np;
    b.iters_add = 1;
    HIPCHK(c, launch_lm(b, K, fp64, KIND_MULTI, f0, c->stream));
  }
  if (f0) {
    F0FlagArgs fa{};
    fa.params = c->d_params.p;
    fa.nvox = c->nvox;
    fa.np = np;
    fa.fb = o.model >= 2 ? 3 : 2;
    fa.K = K;
    fa.list = c->d_list.p;
    fa.count = q + 2;
    fa.refit = c->d_refit.p;
    HIPCHK(c, fdt_launch_f0_flag(&fa, c->stream));
    LmArgs r = a;
    r.list = c->d_list.p;
    r.count = q + 2;
    r.queue = q + 3;
    r.iters_add = 1;
    if (o.model == 1) {
      r.out = c->d_params.p;
      r.out_np = np;
      r.set_f0_slot = 1;
      HIPCHK(c, launch_lm(r, K, fp64, k1, 0, c->stream));
    } else {
      r.out = c->d_s1.p;  // PVM_single_c without f0: 2 + 3K values
      r.out_np = 2 + 3 * K;
      HIPCHK(c, launch_lm(r, K, fp64, KIND_SINGLE_C, 0, c->stream));
      LmArgs m = r;
      m.init = INIT_STAGE1;
      m.in = c->d_s1.p;
      m.in_np = 2 + 3 * K;
      m.G = G;
      m.queue = q + 4;
      m.out = c->d_params.p;
      m.out_np = np;
      m.set_f0_slot = 1;
      HIPCHK(c, launch_lm(m, K, fp64, KIND_MULTI, 0, c->stream));
    }
  }
  if (o.rician) {
    TauArgs t{};
    t.data = c->d_data.p;
    t.N = c->N;
    t.nvox = c->nvox;
    t.acq = c->d_acq.p;
    t.params = c->d_params.p;
    t.np = np;
    t.model = o.model;
    t.f0 = f0;
    t.R = o.R_prior_mean;
    t.refit = f0 ? c->d_refit.p : nullptr;
    t.tau = c->d_tau.p;
    t.acqv = c->gn ? c->d_acqv.p : nullptr;
    t.acqv_stride = c->acqv_stride;
    HIPCHK(c, launch_tau(t, K, fp64, c->stream));
  }
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  c->timed_fit = !c->quiet;
  c->have_params = true; c->have_iters = true;
  c->have_samples = c->have_finish = false;
  return FDT_OK;
}

int fdt_fit(fdt_ctx* c) {
  if (!c) return FDT_EINVAL;
  (void)hipSetDevice(c->device);
  int rc = fit_async(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_lm_stage(fdt_ctx* c, int kind, int K, int f0, int G, float R, const float* start, float* out, int* iters) {
  if (!c || !start || !out) return set_err(c, FDT_EINVAL, "null argument");
  if (kind < 0 || kind > 2 || K < 1 || K > 3) return set_err(c, FDT_EINVAL, "kind=%d K=%d", kind, K);
  if (c->nvox <= 0) return set_err(c, FDT_ESTATE, "no data loaded");
  (void)hipSetDevice(c->device);
  int np = (kind == 2 ? 3 : 2) + 3 * K + (f0 ? 1 : 0);
  size_t n = (size_t)c->nvox * np;
  DevBuf<float> ds, dout;
  DevBuf<int> dit;
  HIPCHK(c, ds.ensure(n));
  HIPCHK(c, dout.ensure(n));
  HIPCHK(c, dit.ensure(c->nvox));
  HIPCHK(c, hipMemcpyAsync(ds.p, start, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  if (int rc = ensure_acqv(c)) return rc;
  HIPCHK(c, c->d_queue.ensure(NQUEUE));
  HIPCHK(c, fdt_launch_zero_ints(c->d_queue.p, NQUEUE, c->stream));
  LmArgs a = base_lm_args(c);
  a.init = INIT_BUF;
  a.in = ds.p;
  a.in_np = np;
  a.G = G;
  a.R = R;
  a.queue = c->d_queue.p;
  a.out = dout.p;
  a.out_np = np;
  a.iters = dit.p;
  HIPCHK(c, launch_lm(a, K, c->o.lm_fp64, kind, f0 ? 1 : 0, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, dout.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (iters) HIPCHK(c, hipMemcpyAsync(iters, dit.p, c->nvox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  ds.release(); dout.release(); dit.release();
  return FDT_OK;
}

int fdt_lm_start(fdt_ctx* c, int kind, int K, int f0, float* out) {
  if (!c || !out) return set_err(c, FDT_EINVAL, "null argument");
  if (kind < 0 || kind > 1 || K < 1 || K > 3) return set_err(c, FDT_EINVAL, "kind=%d K=%d", kind, K);
  if (c->nvox <= 0) return set_err(c, FDT_ESTATE, "no data loaded");
  (void)hipSetDevice(c->device);
  int np = 2 + 3 * K + (f0 ? 1 : 0);
  size_t n = (size_t)c->nvox * np;
  DevBuf<float> dout;
  HIPCHK(c, dout.ensure(n));
  if (int rc = ensure_acqv(c)) return rc;
  HIPCHK(c, c->d_queue.ensure(NQUEUE));
  HIPCHK(c, fdt_launch_zero_ints(c->d_queue.p, NQUEUE, c->stream));
  LmArgs a = base_lm_args(c);
  a.init = INIT_DTI;
  a.lm = 0;
  a.queue = c->d_queue.p;
  a.out = dout.p;
  a.out_np = np;
  HIPCHK(c, launch_lm(a, K, c->o.lm_fp64, kind, f0 ? 1 : 0, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, dout.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dout.release();
  return FDT_OK;
}

int fdt_get_params(fdt_ctx* c, float* params, float* tau) {
  if (!c) return FDT_EINVAL;
  if (!c->have_params) return set_err(c, FDT_ESTATE, "no parameters: call fdt_fit or fdt_set_params");
  (void)hipSetDevice(c->device);
  size_t n = (size_t)c->nvox * np_of(c->o);
  if (params && n) HIPCHK(c, hipMemcpyAsync(params, c->d_params.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (tau && c->nvox) HIPCHK(c, hipMemcpyAsync(tau, c->d_tau.p, c->nvox * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_get_lm_iters(fdt_ctx* c, int* iters) {
  if (!c || !iters) return set_err(c, FDT_EINVAL, "null argument");
  if (!c->have_params || !c->have_iters)
    return set_err(c, FDT_ESTATE, "no LM iteration counts: the current parameters do not come from fdt_fit");
  (void)hipSetDevice(c->device);
  if (c->nvox) HIPCHK(c, hipMemcpyAsync(iters, c->d_iters.p, c->nvox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_set_params(fdt_ctx* c, const float* params, const float* tau) {
  if (!c || !params) return set_err(c, FDT_EINVAL, "null argument");
  if (c->N <= 0) return set_err(c, FDT_ESTATE, "no acquisition");
  (void)hipSetDevice(c->device);
  size_t nv = std::max(c->nvox, 1);
  size_t n = (size_t)c->nvox * np_of(c->o);
  HIPCHK(c, c->d_params.ensure(nv * np_of(c->o)));
  HIPCHK(c, c->d_tau.ensure(nv));
  if (n) HIPCHK(c, hipMemcpyAsync(c->d_params.p, params, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  if (c->nvox) {
    if (tau) HIPCHK(c, hipMemcpyAsync(c->d_tau.p, tau, c->nvox * sizeof(float), hipMemcpyHostToDevice, c->stream));
    else HIPCHK(c, hipMemsetAsync(c->d_tau.p, 0, c->nvox * sizeof(float), c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_params = true; c->have_iters = false;
  c->have_samples = c->have_finish = false;
  return FDT_OK;
}

// buffers, tile and arguments of the chain launch (no launch)
static int mcmc_prepare(fdt_ctx* c, uint64_t vox_offset, McmcArgs* out, mcmc_launch_fn* fn_out) {
  if (!c->have_params) return set_err(c, FDT_ESTATE, "no parameters: call fdt_fit or fdt_set_params");
  const fdt_opts& o = c->o;
  int K = o.nfibres, ns = nsamples_of(o);
  size_t nv = std::max(c->nvox, 1), nsv = (size_t)std::max(ns, 1) * nv;
  HIPCHK(c, c->s_th.ensure(nsv * K));
  HIPCHK(c, c->s_ph.ensure(nsv * K));
  HIPCHK(c, c->s_f.ensure(nsv * K));
  HIPCHK(c, c->s_S0.ensure(nsv));
  HIPCHK(c, c->s_d.ensure(nsv));
  if (o.model >= 2) HIPCHK(c, c->s_dstd.ensure(nsv));
  if (o.model == 3) HIPCHK(c, c->s_R.ensure(nsv));
  if (o.f0) HIPCHK(c, c->s_f0.ensure(nsv));
  if (o.rician) HIPCHK(c, c->s_tau.ensure(nsv));
  HIPCHK(c, c->d_energy.ensure(nv));
  int vpl = 0;
  mcmc_launch_fn fn = nullptr;
  int want = c->vpl_force;
  for (const auto& e : kVpl)
    if (16 * e.vpl >= c->N && e.vpl >= want) { vpl = e.vpl; fn = e.fn; break; }
  if (!fn) return set_err(c, FDT_EUNSUP, "no MCMC tile for N=%d", c->N);
  c->vpl_used = vpl;
  if (int rc = ensure_acqv(c)) return rc;
  McmcArgs a{};
  a.data = c->d_data.p;
  a.params = c->d_params.p;
  a.tau = o.rician ? c->d_tau.p : nullptr;
  a.acq = c->d_acq.p;
  a.shell = c->d_shell.p;
  a.shell_b = c->d_shell_b.p;
  a.nshell = c->nshell;
  a.N = c->N;
  a.Npad = 16 * vpl;
  a.nvox = c->nvox;
  a.np = np_of(o);
  a.model = o.model;
  a.f0 = o.f0;
  a.rician = o.rician;
  a.ardf0 = o.ardf0;
  a.no_ard = o.no_ard;
  a.all_ard = o.all_ard;
  a.fudge = o.fudge;
  a.Rmean = o.R_prior_mean;
  a.Rstd = o.R_prior_std;
  a.Rfudge = o.R_prior_fudge;
  a.nburn = o.nburn;
  a.njumps = o.njumps;
  a.se = o.sampleevery;
  a.upe = o.updateproposalevery;
  a.nsamples = ns;
  a.seed = o.seed;
  a.vox_offset = vox_offset;
  a.th = c->s_th.p; a.ph = c->s_ph.p; a.f = c->s_f.p; a.S0 = c->s_S0.p; a.d = c->s_d.p;
  a.dstd = o.model >= 2 ? c->s_dstd.p : nullptr;
  a.R = o.model == 3 ? c->s_R.p : nullptr;
  a.f0s = o.f0 ? c->s_f0.p : nullptr;
  a.taus = o.rician ? c->s_tau.p : nullptr;
  a.energy = c->d_energy.p;
  a.vstride = c->nvox;
  HIPCHK(c, c->d_widths_out.ensure(nv * 16));
  a.widths_out = c->d_widths_out.p;
  a.widths_in = c->use_widths_in ? c->d_widths_in.p : nullptr;
  if (c->gn) {  // per-voxel b values: no shared b-shell table
    a.acqv = c->d_acqv.p;
    a.acqv_stride = c->acqv_stride;
    a.gdev = c->d_grad.p;
    a.nshell = 0;
  }
  *out = a;
  *fn_out = fn;
  return FDT_OK;
}

static int mcmc_async(fdt_ctx* c, uint64_t vox_offset) {
  McmcArgs a;
  mcmc_launch_fn fn = nullptr;
  int rc = mcmc_prepare(c, vox_offset, &a, &fn);
  if (rc) return rc;
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  HIPCHK(c, fn(&a, c->o.nfibres, c->o.rician, c->stream));
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  c->timed_mcmc = !c->quiet;
  c->have_samples = true;
  c->energy_valid = true;
  c->have_finish = false;
  return FDT_OK;
}

int fdt_mcmc(fdt_ctx* c, uint64_t vox_offset) {
  if (!c) return FDT_EINVAL;
  (void)hipSetDevice(c->device);
  int rc = mcmc_async(c, vox_offset);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

// The voxel loop of the CPU xfibres with neighbour initialisation
// (xfibres.cc:898-921): start, chain and finish of each voxel in mask order,
// each start reading what earlier neighbours left (spatial_stats_kernel).  On
// the device the voxels run in wavefronts t = x + 2y + 4z: a neighbour that
// precedes a voxel in mask order (dz = -1; dz = 0, dy = -1; dz = dy = 0,
// dx = -1) has a smaller t, one that follows it a larger t, so every start
// sees exactly the neighbours the CPU loop has finished, and no two voxels of
// one wavefront are neighbours.  spatial_serial runs one voxel per step
// instead, the CPU's own order.  Three launches per wavefront.
static int spatial_async(fdt_ctx* c, uint64_t vox_offset) {
  if (c->N <= 0) return set_err(c, FDT_ESTATE, "no acquisition");
  if (int rc = cpu_mode_check(c, 2)) return rc;
  const fdt_opts& o = c->o;
  const int nv = c->nvox, K = o.nfibres, ns = nsamples_of(o);
  if (nv > 0 && c->h_gidx.size() != (size_t)nv)
    return set_err(c, FDT_ESTATE, "neighbour initialisation needs the voxel grid: call fdt_set_grid after fdt_load_data");
  if (ns <= 0) return set_err(c, FDT_EINVAL, "neighbour initialisation needs recorded samples (njumps >= sampleevery)");
  size_t nvs = std::max(nv, 1);
  HIPCHK(c, c->d_params.ensure(nvs * np_of(o)));
  HIPCHK(c, c->d_tau.ensure(nvs));
  HIPCHK(c, c->d_been.ensure(nvs));
  HIPCHK(c, c->d_wlist.ensure(nvs));
  HIPCHK(c, c->sp_f.ensure(nvs * K));
  HIPCHK(c, c->sp_d.ensure(nvs));
  HIPCHK(c, c->sp_S0.ensure(nvs));
  HIPCHK(c, c->sp_dyad.ensure(nvs * K * 3));
  HIPCHK(c, fdt_launch_zero_ints(c->d_been.p, nv, c->stream));
  HIPCHK(c, fdt_launch_zero_ints(reinterpret_cast<int*>(c->d_tau.p), nv, c->stream));
  c->have_params = true; c->have_iters = false;
  McmcArgs a;
  mcmc_launch_fn fn = nullptr;
  if (int rc = mcmc_prepare(c, vox_offset, &a, &fn)) return rc;
  // wavefront lists (counting sort by t, mask order inside a wavefront)
  std::vector<int> order(nv), first;
  if (o.spatial_serial) {
    for (int v = 0; v < nv; ++v) order[v] = v;
    for (int v = 0; v <= nv; ++v) first.push_back(v);
  } else if (nv > 0) {
    const long long nx = c->gnx, ny = c->gny;
    const int tmax = (int)((c->gnx - 1) + 2LL * (c->gny - 1) + 4LL * (c->gnz - 1));
    std::vector<int> t(nv), start(tmax + 2, 0);
    for (int v = 0; v < nv; ++v) {
      const long long g = c->h_gidx[v];
      t[v] = (int)(g % nx + 2 * ((g / nx) % ny) + 4 * (g / (nx * ny)));
      ++start[t[v] + 1];
    }
    for (int i = 0; i <= tmax; ++i) start[i + 1] += start[i];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int v = 0; v < nv; ++v) order[fill[t[v]]++] = v;
    for (int i = 0; i <= tmax; ++i)
      if (start[i + 1] > start[i]) first.push_back(start[i]);
    first.push_back(nv);
  }
  // stream-ordered upload (an earlier run on this stream has finished reading the lists before it lands); the
  // host copy lives in the context until the next run, which first waits for this one
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->h_wlist.swap(order);
  if (nv > 0)
    HIPCHK(c, hipMemcpyAsync(c->d_wlist.p, c->h_wlist.data(), (size_t)nv * sizeof(int), hipMemcpyHostToDevice, c->stream));
  CpuStartArgs st = cpu_start_args(c);
  st.vol2mat = c->d_vol2mat.p;
  st.gidx = c->d_gidx.p;
  st.nx = c->gnx; st.ny = c->gny; st.nz = c->gnz;
  st.been = c->d_been.p;
  st.mean_f = c->sp_f.p; st.mean_d = c->sp_d.p; st.mean_S0 = c->sp_S0.p; st.dyad = c->sp_dyad.p;
  SpatialStatsArgs ss{};
  ss.K = K; ss.ns = ns; ss.vstride = nv;
  ss.th = c->s_th.p; ss.ph = c->s_ph.p; ss.f = c->s_f.p; ss.S0 = c->s_S0.p; ss.d = c->s_d.p;
  ss.been = c->d_been.p;
  ss.mean_f = c->sp_f.p; ss.mean_d = c->sp_d.p; ss.mean_S0 = c->sp_S0.p; ss.dyad = c->sp_dyad.p;
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  for (size_t g = 0; g + 1 < first.size(); ++g) {
    const int* L = c->d_wlist.p + first[g];
    const int n = first[g + 1] - first[g];
    st.list = L; st.count = n;
    HIPCHK(c, fdt_launch_cpu_start(&st, c->stream));
    McmcArgs b = a;
    b.list = L;
    b.nvox = n;
    HIPCHK(c, fn(&b, K, 0, c->stream));
    ss.list = L; ss.count = n;
    HIPCHK(c, fdt_launch_spatial_stats(&ss, c->stream));
  }
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  c->timed_fit = false;
  c->timed_mcmc = !c->quiet;
  c->have_samples = true;
  c->energy_valid = true;
  c->have_finish = false;
  c->have_spatial = true;
  return FDT_OK;
}

static int finish_async(fdt_ctx* c) {
  if (!c->have_samples) return set_err(c, FDT_ESTATE, "no samples: call fdt_mcmc first");
  const fdt_opts& o = c->o;
  int K = o.nfibres, ns = nsamples_of(o);
  size_t nv = std::max(c->nvox, 1), nsv = (size_t)std::max(ns, 1) * nv;
  HIPCHK(c, c->o_th.ensure(nsv * K));
  HIPCHK(c, c->o_ph.ensure(nsv * K));
  HIPCHK(c, c->o_f.ensure(nsv * K));
  DevBuf<float>* mb[] = {&c->m_S0, &c->m_d, &c->m_dstd, &c->m_R, &c->m_f0, &c->m_tau};
  for (auto* b : mb) HIPCHK(c, b->ensure(nv));
  HIPCHK(c, c->m_f.ensure(nv * K));
  HIPCHK(c, c->m_dyads.ensure(nv * K * 3));
  HIPCHK(c, c->m_disp.ensure(nv * K));
  FinishArgs a{};
  a.K = K; a.ns = ns; a.nvox = c->nvox; a.model = o.model; a.f0 = o.f0; a.rician = o.rician;
  a.th = c->s_th.p; a.ph = c->s_ph.p; a.f = c->s_f.p; a.S0 = c->s_S0.p; a.d = c->s_d.p;
  a.dstd = o.model >= 2 ? c->s_dstd.p : nullptr;
  a.R = o.model == 3 ? c->s_R.p : nullptr;
  a.f0s = o.f0 ? c->s_f0.p : nullptr;
  a.taus = o.rician ? c->s_tau.p : nullptr;
  a.oth = c->o_th.p; a.oph = c->o_ph.p; a.of = c->o_f.p;
  a.mean_S0 = c->m_S0.p; a.mean_d = c->m_d.p; a.mean_dstd = c->m_dstd.p; a.mean_R = c->m_R.p;
  a.mean_f0 = c->m_f0.p; a.mean_tau = c->m_tau.p; a.mean_f = c->m_f.p; a.dyads = c->m_dyads.p; a.disp = c->m_disp.p;
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
  if (ns > 0) HIPCHK(c, fdt_launch_finish(&a, c->stream));
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
  c->timed_finish = !c->quiet;
  c->have_finish = true;
  return FDT_OK;
}

int fdt_finish(fdt_ctx* c) {
  if (!c) return FDT_EINVAL;
  (void)hipSetDevice(c->device);
  int rc = finish_async(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_get_outputs(fdt_ctx* c, const fdt_outputs* out) {
  if (!c || !out) return set_err(c, FDT_EINVAL, "null argument");
  if (!c->have_finish) return set_err(c, FDT_ESTATE, "no outputs: call fdt_finish first");
  (void)hipSetDevice(c->device);
  const fdt_opts& o = c->o;
  int K = o.nfibres, ns = nsamples_of(o);
  size_t nv = c->nvox, nsv = (size_t)ns * nv;
  auto cp = [&](float* dst, const float* src, size_t n) -> int {
    if (dst && src && n) HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return FDT_OK;
  };
  int rc = 0;
  rc |= cp(out->th, c->o_th.p, nsv * K);
  rc |= cp(out->ph, c->o_ph.p, nsv * K);
  rc |= cp(out->f, c->o_f.p, nsv * K);
  rc |= cp(out->mean_S0, c->m_S0.p, nv);
  rc |= cp(out->mean_d, c->m_d.p, nv);
  if (o.model >= 2) rc |= cp(out->mean_dstd, c->m_dstd.p, nv);
  if (o.model == 3) rc |= cp(out->mean_R, c->m_R.p, nv);
  if (o.f0) rc |= cp(out->mean_f0, c->m_f0.p, nv);
  if (o.rician) rc |= cp(out->mean_tau, c->m_tau.p, nv);
  rc |= cp(out->mean_f, c->m_f.p, nv * K);
  rc |= cp(out->dyads, c->m_dyads.p, nv * K * 3);
  rc |= cp(out->disp, c->m_disp.p, nv * K);
  rc |= cp(out->S0, c->s_S0.p, nsv);
  rc |= cp(out->d, c->s_d.p, nsv);
  if (o.model >= 2) rc |= cp(out->dstd, c->s_dstd.p, nsv);
  if (o.model == 3) rc |= cp(out->R, c->s_R.p, nsv);
  if (o.f0) rc |= cp(out->f0, c->s_f0.p, nsv);
  if (o.rician) rc |= cp(out->tau, c->s_tau.p, nsv);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_run(fdt_ctx* c, uint64_t vox_offset) {
  if (!c) return FDT_EINVAL;
  (void)hipSetDevice(c->device);
  int rc;
  if (cpu_mode(c->o) == 2) {
    rc = spatial_async(c, vox_offset);
  } else {
    rc = fit_async(c);
    if (!rc) rc = mcmc_async(c, vox_offset);
  }
  if (!rc) rc = finish_async(c);
  return rc;
}

int fdt_set_grid(fdt_ctx* c, int nx, int ny, int nz, const long long* gidx) {
  if (!c) return FDT_EINVAL;
  if (nx <= 0 || ny <= 0 || nz <= 0) return set_err(c, FDT_EINVAL, "grid %d x %d x %d", nx, ny, nz);
  const long long nxyz = (long long)nx * ny * nz;
  if (nxyz > (1LL << 31)) return set_err(c, FDT_EINVAL, "grid of %lld voxels", nxyz);
  if (c->nvox > 0 && !gidx) return set_err(c, FDT_EINVAL, "null voxel indices");
  for (int v = 0; v < c->nvox; ++v) {
    if (gidx[v] < 0 || gidx[v] >= nxyz) return set_err(c, FDT_EINVAL, "voxel %d: grid index %lld outside the grid", v, gidx[v]);
    if (v > 0 && gidx[v] <= gidx[v - 1]) return set_err(c, FDT_EINVAL, "voxel %d: grid indices must increase (mask order)", v);
  }
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));  // a run in flight may still read the grid
  std::vector<int> v2m((size_t)nxyz, 0);
  for (int v = 0; v < c->nvox; ++v) v2m[(size_t)gidx[v]] = v + 1;
  HIPCHK(c, c->d_vol2mat.ensure((size_t)nxyz));
  HIPCHK(c, c->d_gidx.ensure(std::max(c->nvox, 1)));
  HIPCHK(c, hipMemcpyAsync(c->d_vol2mat.p, v2m.data(), (size_t)nxyz * sizeof(int), hipMemcpyHostToDevice, c->stream));
  if (c->nvox > 0)
    HIPCHK(c, hipMemcpyAsync(c->d_gidx.p, gidx, (size_t)c->nvox * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->gnx = nx; c->gny = ny; c->gnz = nz;
  c->h_gidx.assign(gidx, gidx + c->nvox);
  return FDT_OK;
}

int fdt_get_spatial_state(fdt_ctx* c, int* been, float* mean_f, float* mean_d, float* mean_S0, double* dyad) {
  if (!c) return FDT_EINVAL;
  if (!c->have_spatial) return set_err(c, FDT_ESTATE, "no neighbour-initialised run");
  (void)hipSetDevice(c->device);
  const size_t nv = c->nvox, K = c->o.nfibres;
  if (nv) {
    if (been) HIPCHK(c, hipMemcpyAsync(been, c->d_been.p, nv * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (mean_f) HIPCHK(c, hipMemcpyAsync(mean_f, c->sp_f.p, nv * K * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (mean_d) HIPCHK(c, hipMemcpyAsync(mean_d, c->sp_d.p, nv * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (mean_S0) HIPCHK(c, hipMemcpyAsync(mean_S0, c->sp_S0.p, nv * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (dyad) HIPCHK(c, hipMemcpyAsync(dyad, c->sp_dyad.p, nv * K * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_sync(fdt_ctx* c) {
  if (!c) return FDT_EINVAL;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_get_timing(fdt_ctx* c, fdt_timing* t) {
  if (!c || !t) return FDT_EINVAL;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::memset(t, 0, sizeof *t);
  if (c->timed_fit) HIPCHK(c, hipEventElapsedTime(&t->ms_fit, c->ev[0], c->ev[1]));
  if (c->timed_mcmc) HIPCHK(c, hipEventElapsedTime(&t->ms_mcmc, c->ev[2], c->ev[3]));
  if (c->timed_finish) HIPCHK(c, hipEventElapsedTime(&t->ms_finish, c->ev[4], c->ev[5]));
  t->mcmc_vpl = c->vpl_used;
  t->mcmc_lanes = 16;
  return FDT_OK;
}

int fdt_set_mcmc_tile(fdt_ctx* c, int vpl) {
  if (!c) return FDT_EINVAL;
  if (vpl < 0) return set_err(c, FDT_EINVAL, "vpl=%d", vpl);
  c->vpl_force = vpl;
  return FDT_OK;
}

// ------------------------------------------------------------------ queue
// A batch of volumes run together: every volume's LM on its own stream
// (forked from and joined to the queue stream), then one persistent chain
// launch over all volumes whose tile the queue kernel carries, then the
// finishing passes.  The first fdt_run_queue runs eagerly (allocating and
// uploading everything) and records the same launches into a hipGraph that
// later runs replay.

}  // extern "C"

struct fdt_queue {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<fdt_ctx*> jobs;
  std::vector<QueueJob> h_jobs;
  std::vector<int> h_first;
  int rows_per_wg = 16;
  DevBuf<QueueJob> d_jobs, d_tune;
  DevBuf<int> d_first, d_counter;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev_fork = nullptr, t0 = nullptr, t1 = nullptr;
  std::string err;
  float ms_last = 0.0f;
  bool ran = false;
};

static int qerr(fdt_queue* q, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (q) q->err = buf;
  return code;
}
#define QHIP(q, expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return qerr(q, FDT_EHIP, "%s: %s", #expr, hipGetErrorString(e_));        \
  } while (0)
#define QCTX(q, c, expr)                                                                           \
  do {                                                                                             \
    int rc_ = (expr);                                                                              \
    if (rc_) return qerr(q, rc_, "%s", (c)->err.c_str());                                          \
  } while (0)

static void queue_invalidate(fdt_queue* q) {
  if (q->exec) (void)hipGraphExecDestroy(q->exec);
  if (q->graph) (void)hipGraphDestroy(q->graph);
  q->exec = nullptr;
  q->graph = nullptr;
}

// jobs the queue kernel carries; Rician, --gradnonlin and the widest tiles run
// their own chain launch inside the batch
static bool in_queue_kernel(const fdt_ctx* c) {
  int vpl = tile_of(c->N, c->vpl_force);
  return !c->o.rician && !c->gn && vpl > 0 && vpl <= QUEUE_MAX_VPL;
}

static int queue_launch_chains(fdt_queue* q, const QueueJob* d_jobs, int njobs, const int* d_first, int total) {
  QueueArgs qa{};
  qa.jobs = d_jobs;
  qa.first_item = d_first;
  qa.njobs = njobs;
  qa.counter = q->d_counter.p;
  QHIP(q, fdt_launch_zero_ints(q->d_counter.p, 1, q->stream));
  QHIP(q, fdt_launch_mcmc_queue(&qa, q->rows_per_wg, total, q->stream));
  return FDT_OK;
}

// issue the whole batch on the queue stream (upload: also rebuild and copy the job table)
static int queue_issue(fdt_queue* q, bool upload) {
  QHIP(q, hipEventRecord(q->ev_fork, q->stream));
  for (fdt_ctx* c : q->jobs) {
    QHIP(q, hipStreamWaitEvent(c->stream, q->ev_fork, 0));
    QCTX(q, c, fit_async(c));
    if (!in_queue_kernel(c)) QCTX(q, c, mcmc_async(c, c->vox_offset));
    QHIP(q, hipEventRecord(c->ev_done, c->stream));
    QHIP(q, hipStreamWaitEvent(q->stream, c->ev_done, 0));
  }
  if (upload) {
    q->h_jobs.clear();
    q->h_first.assign(1, 0);
    // items of the costliest volumes first (cost per voxel ~ N * (K + 1)), so
    // that the launch ends on short items
    std::vector<fdt_ctx*> order(q->jobs.begin(), q->jobs.end());
    std::stable_sort(order.begin(), order.end(), [](const fdt_ctx* x, const fdt_ctx* y) {
      return (long)x->N * (x->o.nfibres + 1) > (long)y->N * (y->o.nfibres + 1);
    });
    for (fdt_ctx* c : order) {
      if (!in_queue_kernel(c)) continue;
      QueueJob j{};
      mcmc_launch_fn fn = nullptr;
      QCTX(q, c, mcmc_prepare(c, c->vox_offset, &j.a, &fn));
      j.K = c->o.nfibres;
      j.vpl = c->vpl_used;
      q->h_jobs.push_back(j);
      q->h_first.push_back(q->h_first.back() + (c->nvox + q->rows_per_wg - 1) / q->rows_per_wg);
    }
    if (!q->h_jobs.empty()) {
      QHIP(q, q->d_jobs.ensure(q->h_jobs.size()));
      QHIP(q, q->d_first.ensure(q->h_first.size()));
      QHIP(q, hipMemcpyAsync(q->d_jobs.p, q->h_jobs.data(), sizeof(QueueJob) * q->h_jobs.size(), hipMemcpyHostToDevice,
                             q->stream));
      QHIP(q, hipMemcpyAsync(q->d_first.p, q->h_first.data(), sizeof(int) * q->h_first.size(), hipMemcpyHostToDevice,
                             q->stream));
    }
  }
  if (!q->h_jobs.empty())
    if (int rc = queue_launch_chains(q, q->d_jobs.p, (int)q->h_jobs.size(), q->d_first.p, q->h_first.back())) return rc;
  QHIP(q, hipEventRecord(q->ev_fork, q->stream));
  for (fdt_ctx* c : q->jobs) {
    c->have_samples = true;
    c->energy_valid = false;
    QHIP(q, hipStreamWaitEvent(c->stream, q->ev_fork, 0));
    QCTX(q, c, finish_async(c));
    QHIP(q, hipEventRecord(c->ev_done, c->stream));
    QHIP(q, hipStreamWaitEvent(q->stream, c->ev_done, 0));
  }
  return FDT_OK;
}

extern "C" {

fdt_queue* fdt_queue_create(int device) {
  fdt_queue* q = new fdt_queue;
  q->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&q->stream) != hipSuccess ||
      hipEventCreateWithFlags(&q->ev_fork, hipEventDisableTiming) != hipSuccess || hipEventCreate(&q->t0) != hipSuccess ||
      hipEventCreate(&q->t1) != hipSuccess || q->d_counter.ensure(1) != hipSuccess) {
    delete q;
    return nullptr;
  }
  return q;
}

void fdt_queue_destroy(fdt_queue* q) {
  if (!q) return;
  (void)hipSetDevice(q->device);
  if (q->stream) (void)hipStreamSynchronize(q->stream);
  queue_invalidate(q);
  for (fdt_ctx* c : q->jobs) fdt_destroy(c);
  q->d_jobs.release();
  q->d_tune.release();
  q->d_first.release();
  q->d_counter.release();
  if (q->ev_fork) (void)hipEventDestroy(q->ev_fork);
  if (q->t0) (void)hipEventDestroy(q->t0);
  if (q->t1) (void)hipEventDestroy(q->t1);
  if (q->stream) (void)hipStreamDestroy(q->stream);
  delete q;
}

const char* fdt_queue_last_error(const fdt_queue* q) { return q ? q->err.c_str() : "null queue"; }

int fdt_enqueue_batch(fdt_queue* q, const fdt_opts* o, int N, const float* bvals, const float* bvecs, const float* data,
                      int layout, int nvox, uint64_t vox_offset) {
  if (!q || !o) return FDT_EINVAL;
  (void)hipSetDevice(q->device);
  if (cpu_mode(*o) == 2)
    return qerr(q, FDT_EUNSUP, "job %zu: neighbour initialisation runs voxel wavefronts one after another and has no "
                               "place in a batch; use fdt_run, or nospat for the tensor start", q->jobs.size());
  fdt_ctx* c = fdt_create(q->device, o);
  if (!c) return qerr(q, FDT_EINVAL, "job %zu: invalid options or device", q->jobs.size());
  int rc = fdt_set_acquisition(c, N, bvals, bvecs);
  if (!rc) rc = fdt_load_data(c, data, layout, nvox);
  if (!rc && hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming) != hipSuccess) rc = FDT_EHIP;
  if (rc) {
    qerr(q, rc, "job %zu: %s", q->jobs.size(), c->err.c_str());
    fdt_destroy(c);
    return rc;
  }
  c->vox_offset = vox_offset;
  c->quiet = true;
  q->jobs.push_back(c);
  queue_invalidate(q);
  q->ran = false;
  return (int)q->jobs.size() - 1;
}

int fdt_queue_set_grad_dev(fdt_queue* q, int job, const float* grad, int layout) {
  if (!q) return FDT_EINVAL;
  if (job < 0 || job >= (int)q->jobs.size()) return qerr(q, FDT_EINVAL, "job %d out of range", job);
  (void)hipSetDevice(q->device);
  fdt_ctx* c = q->jobs[job];
  if (int rc = fdt_load_grad_dev(c, grad, layout, c->nvox)) return qerr(q, rc, "job %d: %s", job, c->err.c_str());
  queue_invalidate(q);
  q->ran = false;
  return FDT_OK;
}

int fdt_queue_set_tiles(fdt_queue* q, int rows_per_wg, int job, int vpl) {
  if (!q) return FDT_EINVAL;
  if (rows_per_wg) {
    if (rows_per_wg != 4 && rows_per_wg != 8 && rows_per_wg != 16)
      return qerr(q, FDT_EINVAL, "rows_per_wg=%d: must be 4, 8 or 16", rows_per_wg);
    q->rows_per_wg = rows_per_wg;
  }
  if (job >= 0) {
    if (job >= (int)q->jobs.size()) return qerr(q, FDT_EINVAL, "job %d out of range", job);
    if (vpl < 0 || (vpl && !tile_of(q->jobs[job]->N, vpl))) return qerr(q, FDT_EINVAL, "vpl=%d", vpl);
    q->jobs[job]->vpl_force = vpl;
  }
  queue_invalidate(q);
  q->ran = false;
  return FDT_OK;
}

int fdt_run_queue(fdt_queue* q) {
  if (!q) return FDT_EINVAL;
  if (q->jobs.empty()) return qerr(q, FDT_ESTATE, "empty queue");
  (void)hipSetDevice(q->device);
  QHIP(q, hipEventRecord(q->t0, q->stream));
  if (!q->exec) {
    if (int rc = queue_issue(q, true)) return rc;
    QHIP(q, hipEventRecord(q->t1, q->stream));
    QHIP(q, hipStreamSynchronize(q->stream));
    QHIP(q, hipEventElapsedTime(&q->ms_last, q->t0, q->t1));
    // record the same launches into a graph for the next runs
    QHIP(q, hipStreamBeginCapture(q->stream, hipStreamCaptureModeRelaxed));
    int rc = queue_issue(q, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(q->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    QHIP(q, e);
    q->graph = g;
    QHIP(q, hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0));
  } else {
    QHIP(q, hipGraphLaunch(q->exec, q->stream));
    QHIP(q, hipEventRecord(q->t1, q->stream));
    QHIP(q, hipStreamSynchronize(q->stream));
    QHIP(q, hipEventElapsedTime(&q->ms_last, q->t0, q->t1));
  }
  q->ran = true;
  return FDT_OK;
}

int fdt_queue_num_jobs(const fdt_queue* q) { return q ? (int)q->jobs.size() : FDT_EINVAL; }

int fdt_queue_get_outputs(fdt_queue* q, int job, const fdt_outputs* out) {
  if (!q || !out) return FDT_EINVAL;
  if (job < 0 || job >= (int)q->jobs.size()) return qerr(q, FDT_EINVAL, "job %d out of range", job);
  if (!q->ran) return qerr(q, FDT_ESTATE, "fdt_run_queue has not run since the last change");
  QHIP(q, hipStreamSynchronize(q->stream));
  QCTX(q, q->jobs[job], fdt_get_outputs(q->jobs[job], out));
  return FDT_OK;
}

int fdt_queue_get_params(fdt_queue* q, int job, float* params, float* tau) {
  if (!q) return FDT_EINVAL;
  if (job < 0 || job >= (int)q->jobs.size()) return qerr(q, FDT_EINVAL, "job %d out of range", job);
  QHIP(q, hipStreamSynchronize(q->stream));
  QCTX(q, q->jobs[job], fdt_get_params(q->jobs[job], params, tau));
  return FDT_OK;
}

int fdt_queue_timing(fdt_queue* q, float* ms_last, int* rows_per_wg, int* vpl_of_job, int njobs) {
  if (!q) return FDT_EINVAL;
  if (ms_last) *ms_last = q->ms_last;
  if (rows_per_wg) *rows_per_wg = q->rows_per_wg;
  if (vpl_of_job)
    for (int j = 0; j < njobs && j < (int)q->jobs.size(); ++j) vpl_of_job[j] = tile_of(q->jobs[j]->N, q->jobs[j]->vpl_force);
  return FDT_OK;
}

// Autotuner: times the queue kernel on a sample of every volume (first
// `sample_vox` voxels, `jumps` sweeps, no burn-in) for each voxels-per-
// workgroup choice (4, 8, 16), keeps the fastest, then for each volume tries
// the next larger volumes-per-lane tile and keeps it when faster.  Needs LM
// parameters: the first call runs the batch once.
int fdt_queue_autotune(fdt_queue* q, int jumps, int sample_vox, float* ms_best) {
  if (!q) return FDT_EINVAL;
  if (q->jobs.empty()) return qerr(q, FDT_ESTATE, "empty queue");
  if (jumps <= 0 || sample_vox <= 0) return qerr(q, FDT_EINVAL, "jumps=%d sample_vox=%d", jumps, sample_vox);
  (void)hipSetDevice(q->device);
  if (!q->ran)
    if (int rc = fdt_run_queue(q)) return rc;
  auto time_config = [&](float* ms) -> int {
    std::vector<QueueJob> tj;
    std::vector<int> first(1, 0);
    std::vector<fdt_ctx*> order(q->jobs.begin(), q->jobs.end());
    std::stable_sort(order.begin(), order.end(), [](const fdt_ctx* x, const fdt_ctx* y) {
      return (long)x->N * (x->o.nfibres + 1) > (long)y->N * (y->o.nfibres + 1);
    });
    for (fdt_ctx* c : order) {
      if (!in_queue_kernel(c)) continue;
      QueueJob j{};
      mcmc_launch_fn fn = nullptr;
      QCTX(q, c, mcmc_prepare(c, c->vox_offset, &j.a, &fn));
      j.K = c->o.nfibres;
      j.vpl = c->vpl_used;
      j.a.nvox = std::min(c->nvox, sample_vox);
      j.a.nburn = 0;
      j.a.njumps = jumps;
      j.a.se = jumps;
      j.a.nsamples = 1;  // writes sample 0 of the volume's buffers (stride nvox kept)
      j.a.energy = nullptr;
      tj.push_back(j);
      first.push_back(first.back() + (j.a.nvox + q->rows_per_wg - 1) / q->rows_per_wg);
    }
    if (tj.empty()) { *ms = 0.0f; return FDT_OK; }
    QHIP(q, q->d_tune.ensure(tj.size()));
    DevBuf<int> dfirst;
    QHIP(q, dfirst.ensure(first.size()));
    QHIP(q, hipMemcpyAsync(q->d_tune.p, tj.data(), sizeof(QueueJob) * tj.size(), hipMemcpyHostToDevice, q->stream));
    QHIP(q, hipMemcpyAsync(dfirst.p, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice, q->stream));
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      QHIP(q, hipEventRecord(q->t0, q->stream));
      if (int rc = queue_launch_chains(q, q->d_tune.p, (int)tj.size(), dfirst.p, first.back())) return rc;
      QHIP(q, hipEventRecord(q->t1, q->stream));
      QHIP(q, hipStreamSynchronize(q->stream));
      float ms = 0.0f;
      QHIP(q, hipEventElapsedTime(&ms, q->t0, q->t1));
      if (rep > 0) best = std::min(best, ms);
    }
    dfirst.release();
    *ms = best;
    return FDT_OK;
  };
  float best = 1e30f;
  int best_rows = q->rows_per_wg;
  for (int rows : {4, 8, 16}) {
    q->rows_per_wg = rows;
    float ms;
    if (int rc = time_config(&ms)) return rc;
    if (ms < best) { best = ms; best_rows = rows; }
  }
  q->rows_per_wg = best_rows;
  for (fdt_ctx* c : q->jobs) {
    if (!in_queue_kernel(c)) continue;
    int cur = tile_of(c->N, c->vpl_force);
    int next = tile_of(c->N, cur + 1);
    if (!next || next > QUEUE_MAX_VPL) continue;
    int keep = c->vpl_force;
    c->vpl_force = next;
    float ms;
    if (int rc = time_config(&ms)) return rc;
    if (ms < best) best = ms;
    else c->vpl_force = keep;
  }
  if (ms_best) *ms_best = best;
  queue_invalidate(q);
  q->ran = false;  // the timing runs overwrote sample 0 of each volume
  return FDT_OK;
}

}  // extern "C"

// -------------------------------------------------------------- surrogate
// The learned initialiser: an MLP N -> 512 -> 512 -> 256 -> 28 (GELU between
// layers, BF16 weights and activations, FP32 accumulation, FP32 master
// weights and Adam moments).  Flat weight layout: W1 [N][512], b1, W2
// [512][512], b2, W3 [512][256], b3, W4 [256][28], b4 (W[l] is in x out).

struct fdt_surrogate {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_in = 0;
  int dims[SG_LAYERS + 1] = {};
  size_t woff[SG_LAYERS] = {}, boff[SG_LAYERS] = {}, nw = 0;
  DevBuf<float> w, m, v, g;
  DevBuf<uint16_t> wb;
  int step = 0;
  DevBuf<uint16_t> X;  // normalised training signals [n][n_in]
  DevBuf<float> T;     // targets [n][SG_OUT]
  DevBuf<float> bvals;
  int ntrain = 0;
  int cap = 0;  // batch capacity of the activation buffers
  DevBuf<float> Z[SG_LAYERS];
  DevBuf<uint16_t> A[SG_LAYERS];
  DevBuf<float> dA;
  DevBuf<uint16_t> dZ, Xb;
  DevBuf<float> loss_part, S0;
  std::string err;
};

static int serr(fdt_surrogate* s, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (s) s->err = buf;
  return code;
}
#define SHIP(s, expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return serr(s, FDT_EHIP, "%s: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

static int sg_reserve(fdt_surrogate* s, int b) {
  if (b <= s->cap) return FDT_OK;
  for (int l = 0; l < SG_LAYERS; ++l) {
    SHIP(s, s->Z[l].ensure((size_t)b * s->dims[l + 1]));
    SHIP(s, s->A[l].ensure((size_t)b * s->dims[l + 1]));
  }
  SHIP(s, s->dA.ensure((size_t)b * 512));
  SHIP(s, s->dZ.ensure((size_t)b * 512));
  SHIP(s, s->Xb.ensure((size_t)b * s->n_in));
  SHIP(s, s->S0.ensure((size_t)b));
  SHIP(s, s->loss_part.ensure((size_t)(b + 255) / 256));
  s->cap = b;
  return FDT_OK;
}

static GemmArgs gemm(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.sam = sam; g.sak = sak;
  g.B = B; g.sbk = sbk; g.sbn = sbn;
  g.ldc = N;
  return g;
}

// forward of b rows of X (bf16 [b][n_in]); keeps Z (pre-activations) when
// `keep`, the output in Z[3]
static int sg_forward(fdt_surrogate* s, const uint16_t* X, int b, bool keep) {
  const uint16_t* in = X;
  for (int l = 0; l < SG_LAYERS; ++l) {
    const int K = s->dims[l], N = s->dims[l + 1];
    GemmArgs g = gemm(b, N, K, in, K, 1, s->wb.p + s->woff[l], N, 1);
    g.bias = s->w.p + s->boff[l];
    const bool last = l == SG_LAYERS - 1;
    g.C = (keep || last) ? s->Z[l].p : nullptr;
    g.Cb = last ? nullptr : s->A[l].p;
    g.act = last ? 0 : 1;
    SHIP(s, fdt_launch_gemm_bf16(&g, s->stream));
    in = s->A[l].p;
  }
  return FDT_OK;
}

// gradients of the mean loss over b rows into s->g; loss sum into *loss (host)
static int sg_backward(fdt_surrogate* s, const uint16_t* X, const float* T, int b, float* loss) {
  const int nb = (b + 255) / 256;
  SHIP(s, fdt_launch_sg_loss(s->Z[SG_LAYERS - 1].p, T, b, s->dZ.p, s->loss_part.p, 0.1f, s->stream));
  for (int l = SG_LAYERS - 1; l >= 0; --l) {
    const int K = s->dims[l], N = s->dims[l + 1];
    const uint16_t* in = l ? s->A[l - 1].p : X;
    // dW = in^T . dZ  (K x N, reduced over the batch)
    GemmArgs gw = gemm(K, N, b, in, 1, K, s->dZ.p, N, 1);
    gw.C = s->g.p + s->woff[l];
    SHIP(s, fdt_launch_gemm_bf16(&gw, s->stream));
    SHIP(s, fdt_launch_mlp_ew(1, s->dZ.p, nullptr, s->g.p + s->boff[l], 0, b, N, s->stream));
    if (l == 0) break;
    // dA = dZ . W^T  (b x K), then dZ = dA * gelu'(Z[l-1])
    GemmArgs ga = gemm(b, K, N, s->dZ.p, N, 1, s->wb.p + s->woff[l], 1, N);
    ga.C = s->dA.p;
    SHIP(s, fdt_launch_gemm_bf16(&ga, s->stream));
    SHIP(s, fdt_launch_mlp_ew(0, s->dA.p, s->Z[l - 1].p, s->dZ.p, (size_t)b * K, 0, 0, s->stream));
  }
  if (loss) {
    std::vector<float> part(nb);
    SHIP(s, hipMemcpyAsync(part.data(), s->loss_part.p, sizeof(float) * nb, hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    double t = 0.0;
    for (float x : part) t += x;
    *loss = (float)(t / b);
  }
  return FDT_OK;
}

extern "C" {

fdt_surrogate* fdt_surrogate_create(int device, int n_in, uint64_t seed) {
  if (n_in < 7 || n_in > MAX_N) return nullptr;
  fdt_surrogate* s = new fdt_surrogate;
  s->device = device;
  s->n_in = n_in;
  const int d[SG_LAYERS + 1] = {n_in, SG_HID0, SG_HID1, SG_HID2, SG_OUT};
  size_t off = 0;
  for (int l = 0; l <= SG_LAYERS; ++l) s->dims[l] = d[l];
  for (int l = 0; l < SG_LAYERS; ++l) {
    s->woff[l] = off;
    off += (size_t)d[l] * d[l + 1];
    s->boff[l] = off;
    off += d[l + 1];
  }
  s->nw = off;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess || s->w.ensure(off) ||
      s->m.ensure(off) || s->v.ensure(off) || s->g.ensure(off) || s->wb.ensure(off)) {
    fdt_surrogate_destroy(s);
    return nullptr;
  }
  // He-normal weights (std sqrt(2/fan_in)) for the GELU layers, sqrt(1/fan_in) for the output, zero biases
  std::vector<float> h(off, 0.0f);
  std::mt19937_64 gen(seed);
  for (int l = 0; l < SG_LAYERS; ++l) {
    std::normal_distribution<float> nd(0.0f, std::sqrt((l == SG_LAYERS - 1 ? 1.0f : 2.0f) / d[l]));
    for (size_t i = 0; i < (size_t)d[l] * d[l + 1]; ++i) h[s->woff[l] + i] = nd(gen);
  }
  if (fdt_surrogate_load(s, h.data(), off) != FDT_OK) {
    fdt_surrogate_destroy(s);
    return nullptr;
  }
  return s;
}

void fdt_surrogate_destroy(fdt_surrogate* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  DevBuf<float>* fb[] = {&s->w, &s->m, &s->v, &s->g, &s->T, &s->bvals, &s->dA, &s->loss_part, &s->S0};
  for (auto* b : fb) b->release();
  s->wb.release(); s->X.release(); s->dZ.release(); s->Xb.release();
  for (int l = 0; l < SG_LAYERS; ++l) { s->Z[l].release(); s->A[l].release(); }
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char* fdt_surrogate_last_error(const fdt_surrogate* s) { return s ? s->err.c_str() : "null surrogate"; }

size_t fdt_surrogate_num_weights(const fdt_surrogate* s) { return s ? s->nw : 0; }

int fdt_surrogate_load(fdt_surrogate* s, const float* w, size_t n) {
  if (!s || !w) return FDT_EINVAL;
  if (n != s->nw) return serr(s, FDT_EINVAL, "%zu weights given, the network has %zu", n, s->nw);
  (void)hipSetDevice(s->device);
  SHIP(s, hipMemcpyAsync(s->w.p, w, sizeof(float) * n, hipMemcpyHostToDevice, s->stream));
  SHIP(s, fdt_launch_mlp_ew(2, s->w.p, nullptr, s->wb.p, n, 0, 0, s->stream));
  SHIP(s, hipMemsetAsync(s->m.p, 0, sizeof(float) * n, s->stream));
  SHIP(s, hipMemsetAsync(s->v.p, 0, sizeof(float) * n, s->stream));
  SHIP(s, hipStreamSynchronize(s->stream));
  s->step = 0;
  return FDT_OK;
}

int fdt_surrogate_save(fdt_surrogate* s, float* w, size_t n) {
  if (!s || !w) return FDT_EINVAL;
  if (n != s->nw) return serr(s, FDT_EINVAL, "buffer of %zu, the network has %zu weights", n, s->nw);
  (void)hipSetDevice(s->device);
  SHIP(s, hipMemcpyAsync(w, s->w.p, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
  SHIP(s, hipStreamSynchronize(s->stream));
  return FDT_OK;
}

int fdt_surrogate_set_training_data(fdt_surrogate* s, const float* signals, const float* bvals, const float* targets,
                                    int n) {
  if (!s || !signals || !bvals || !targets || n <= 0) return FDT_EINVAL;
  (void)hipSetDevice(s->device);
  const int N = s->n_in;
  DevBuf<float> raw;
  SHIP(s, raw.ensure((size_t)n * N));
  SHIP(s, s->bvals.ensure(N));
  SHIP(s, s->X.ensure((size_t)n * N));
  SHIP(s, s->T.ensure((size_t)n * SG_OUT));
  SHIP(s, hipMemcpyAsync(raw.p, signals, sizeof(float) * n * N, hipMemcpyHostToDevice, s->stream));
  SHIP(s, hipMemcpyAsync(s->bvals.p, bvals, sizeof(float) * N, hipMemcpyHostToDevice, s->stream));
  SHIP(s, hipMemcpyAsync(s->T.p, targets, sizeof(float) * n * SG_OUT, hipMemcpyHostToDevice, s->stream));
  SHIP(s, fdt_launch_normalize(raw.p, N, n, s->bvals.p, 50.0f, s->X.p, nullptr, s->stream));
  SHIP(s, hipStreamSynchronize(s->stream));
  raw.release();
  s->ntrain = n;
  return FDT_OK;
}

// gradients of the mean loss over training rows [first, first + batch) (no update)
int fdt_surrogate_gradients(fdt_surrogate* s, int first, int batch, float* grads, float* loss) {
  if (!s || batch <= 0 || first < 0 || first + batch > s->ntrain) return serr(s, FDT_EINVAL, "bad batch");
  (void)hipSetDevice(s->device);
  if (int rc = sg_reserve(s, batch)) return rc;
  const uint16_t* X = s->X.p + (size_t)first * s->n_in;
  if (int rc = sg_forward(s, X, batch, true)) return rc;
  if (int rc = sg_backward(s, X, s->T.p + (size_t)first * SG_OUT, batch, loss)) return rc;
  if (grads) {
    SHIP(s, hipMemcpyAsync(grads, s->g.p, sizeof(float) * s->nw, hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
  }
  return FDT_OK;
}

// one Adam step on training rows [first, first + batch)
int fdt_surrogate_train_step(fdt_surrogate* s, int first, int batch, float lr, float* loss) {
  if (int rc = fdt_surrogate_gradients(s, first, batch, nullptr, loss)) return rc;
  ++s->step;
  SHIP(s, fdt_launch_adam(s->w.p, s->m.p, s->v.p, s->g.p, s->wb.p, s->nw, lr, 0.9f, 0.999f, 1e-8f, s->step, s->stream));
  if (!loss) SHIP(s, hipStreamSynchronize(s->stream));
  return FDT_OK;
}

// network outputs for host signals [n][n_in] (tests, evaluation)
int fdt_surrogate_forward(fdt_surrogate* s, const float* signals, const float* bvals, int n, float* out) {
  if (!s || !signals || !bvals || !out || n <= 0) return FDT_EINVAL;
  (void)hipSetDevice(s->device);
  const int N = s->n_in;
  DevBuf<float> raw;
  SHIP(s, raw.ensure((size_t)n * N));
  SHIP(s, s->bvals.ensure(N));
  SHIP(s, hipMemcpyAsync(raw.p, signals, sizeof(float) * n * N, hipMemcpyHostToDevice, s->stream));
  SHIP(s, hipMemcpyAsync(s->bvals.p, bvals, sizeof(float) * N, hipMemcpyHostToDevice, s->stream));
  const int chunk = 65536;
  if (int rc = sg_reserve(s, std::min(n, chunk))) return rc;
  for (int c0 = 0; c0 < n; c0 += chunk) {
    const int b = std::min(chunk, n - c0);
    SHIP(s, fdt_launch_normalize(raw.p + (size_t)c0 * N, N, b, s->bvals.p, 50.0f, s->Xb.p, nullptr, s->stream));
    if (int rc = sg_forward(s, s->Xb.p, b, false)) return rc;
    SHIP(s, hipMemcpyAsync(out + (size_t)c0 * SG_OUT, s->Z[SG_LAYERS - 1].p, sizeof(float) * b * SG_OUT,
                           hipMemcpyDeviceToHost, s->stream));
  }
  SHIP(s, hipStreamSynchronize(s->stream));
  raw.release();
  return FDT_OK;
}

}  // extern "C"

// Network start values and proposal widths for the context's resident voxels
// (chunks of 64k); also the predicted direction of fibre 1 per voxel.
static int sg_infer(fdt_surrogate* s, fdt_ctx* c, DevBuf<float>& dir) {
  const fdt_opts& o = c->o;
  const int np = np_of(o), nv = c->nvox;
  if (s->n_in != c->N) return serr(s, FDT_EINVAL, "network input %d, acquisition %d", s->n_in, c->N);
  SHIP(s, c->d_params.ensure((size_t)std::max(nv, 1) * np));
  SHIP(s, c->d_widths_in.ensure((size_t)std::max(nv, 1) * 16));
  SHIP(s, dir.ensure((size_t)std::max(nv, 1) * 3));
  SHIP(s, c->d_tau.ensure(std::max(nv, 1)));
  const int chunk = 65536;
  if (int rc = sg_reserve(s, std::min(std::max(nv, 1), chunk))) return rc;
  for (int c0 = 0; c0 < nv; c0 += chunk) {
    const int b = std::min(chunk, nv - c0);
    SHIP(s, fdt_launch_normalize(c->d_data.p + (size_t)c0 * c->N, c->N, b, c->d_bvals.p, 50.0f, s->Xb.p, s->S0.p,
                                 c->stream));
    // the network runs on the context's stream
    hipStream_t keep = s->stream;
    s->stream = c->stream;
    int rc = sg_forward(s, s->Xb.p, b, false);
    s->stream = keep;
    if (rc) return rc;
    SgDecodeArgs d{};
    d.O = s->Z[SG_LAYERS - 1].p;
    d.S0 = s->S0.p;
    d.nvox = b;
    d.model = o.model;
    d.K = o.nfibres;
    d.f0 = o.f0;
    d.np = np;
    d.params = c->d_params.p + (size_t)c0 * np;
    d.widths = c->d_widths_in.p + (size_t)c0 * 16;
    d.dir = dir.p + (size_t)c0 * 3;
    SHIP(s, fdt_launch_sg_decode(&d, c->stream));
  }
  if (o.rician) {
    // Rician precision of the learned start: 1/var of its residuals, as after the LM
    TauArgs t{};
    t.data = c->d_data.p;
    t.N = c->N;
    t.nvox = nv;
    t.acq = c->d_acq.p;
    t.params = c->d_params.p;
    t.np = np;
    t.model = o.model;
    t.f0 = o.f0;
    t.R = o.R_prior_mean;
    t.tau = c->d_tau.p;
    SHIP(s, launch_tau(t, o.nfibres, 0, c->stream));
  }
  return FDT_OK;
}

extern "C" {

int fdt_surrogate_infer(fdt_surrogate* s, fdt_ctx* c) {
  if (!s || !c) return FDT_EINVAL;
  (void)hipSetDevice(c->device);
  DevBuf<float> dir;
  if (int rc = sg_infer(s, c, dir)) return rc;
  SHIP(s, hipStreamSynchronize(c->stream));
  dir.release();
  c->have_params = true; c->have_iters = false;
  c->use_widths_in = true;
  c->have_samples = c->have_finish = false;
  return FDT_OK;
}

// Learned start + short burn-in, with the plain path as fallback: the network
// gives start values and proposal widths, the chains run nburn_short burn-in
// sweeps and the recorded ones; voxels whose final sum of squares exceeds
// chi2_max * N * sigma^2 (sum of squares of the final state's residuals y - S0 M
// for either noise model; sigma^2: variance of the b=0 volumes) or whose
// posterior mean dyad of fibre 1 lies more than ang_max_deg from the predicted
// direction are fitted with the plain LM and rerun with the configured
// burn-in.  *n_fallback receives the number of rerun voxels.
int fdt_run_surrogate(fdt_ctx* c, fdt_surrogate* s, uint64_t vox_offset, int nburn_short, float chi2_max,
                      float ang_max_deg, int* n_fallback) {
  if (!c || !s) return FDT_EINVAL;
  if (nburn_short < 0) return set_err(c, FDT_EINVAL, "nburn_short=%d", nburn_short);
  if (c->gn) return set_err(c, FDT_EUNSUP, "the learned start has no per-voxel acquisitions (--gradnonlin)");
  (void)hipSetDevice(c->device);
  const fdt_opts& o = c->o;
  const int nv = c->nvox, K = o.nfibres;
  DevBuf<float> dir;
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  if (int rc = sg_infer(s, c, dir)) return set_err(c, rc, "%s", s->err.c_str());
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  c->timed_fit = true;
  c->have_params = true; c->have_iters = false;
  // short chains from the learned start
  c->use_widths_in = true;
  McmcArgs a;
  mcmc_launch_fn fn = nullptr;
  if (int rc = mcmc_prepare(c, vox_offset, &a, &fn)) return rc;
  a.nburn = nburn_short;
  a.energy = nullptr;
  a.ssr = c->d_energy.p;  // the sum-of-squares test below, for either noise model
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  HIPCHK(c, fn(&a, K, o.rician, c->stream));
  c->have_samples = true;
  c->energy_valid = false;  // d_energy holds the sum-of-squares test's input here
  if (int rc = finish_async(c)) return rc;
  // fallback list
  HIPCHK(c, c->d_queue.ensure(NQUEUE));
  HIPCHK(c, c->d_list.ensure(std::max(nv, 1)));
  HIPCHK(c, fdt_launch_zero_ints(c->d_queue.p, NQUEUE, c->stream));
  SgFallbackArgs fa{};
  fa.data = c->d_data.p;
  fa.N = c->N;
  fa.nvox = nv;
  fa.bvals = c->d_bvals.p;
  fa.b0_max = 50.0f;
  fa.energy = c->d_energy.p;
  fa.dyads = c->m_dyads.p;
  fa.K = K;
  HIPCHK(c, c->d_refit.ensure(std::max(nv, 1)));
  fa.flag = c->d_refit.p;
  fa.dir = dir.p;
  // sum of squares of the short chain's final residuals against chi2_max N sigma^2 (for Rician data
  // the residual of the mean signal, which the noise floor inflates little at the SNR of diffusion MRI)
  fa.chi2_max = chi2_max;
  fa.cos_min = std::cos(ang_max_deg * 3.14159265358979f / 180.0f);
  fa.list = c->d_list.p;
  fa.count = c->d_queue.p + 7;
  HIPCHK(c, fdt_launch_sg_fallback(&fa, c->stream));
  int nfb = 0;
  HIPCHK(c, hipMemcpyAsync(&nfb, c->d_queue.p + 7, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n_fallback) *n_fallback = nfb;
  if (nfb > 0) {
    // plain LM on the listed voxels (stage 1 from the tensor, then PVM_multi for model >= 2)
    const int f0 = o.f0 ? 1 : 0;
    const int np = np_of(o), np1 = 2 + 3 * K + f0;
    int* q = c->d_queue.p;
    HIPCHK(c, fdt_launch_zero_ints(q, 4, c->stream));
    LmArgs l = base_lm_args(c);
    l.rician = o.rician;
    l.list = c->d_list.p;
    l.count = c->d_queue.p + 7;
    l.init = INIT_DTI;
    l.queue = q;
    const int k1 = o.model == 1 ? (o.nonlin ? KIND_SINGLE : KIND_SINGLE_C) : KIND_SINGLE;
    if (o.model == 1) {
      l.out = c->d_params.p;
      l.out_np = np;
      HIPCHK(c, launch_lm(l, K, o.lm_fp64, k1, f0, c->stream));
    } else {
      HIPCHK(c, c->d_s1.ensure((size_t)nv * np1));
      l.out = c->d_s1.p;
      l.out_np = np1;
      HIPCHK(c, launch_lm(l, K, o.lm_fp64, k1, f0, c->stream));
      LmArgs m = l;
      m.init = INIT_STAGE1;
      m.in = c->d_s1.p;
      m.in_np = np1;
      m.G = o.model == 3 ? 2 : 0;
      m.queue = q + 1;
      m.out = c->d_params.p;
      m.out_np = np;
      HIPCHK(c, launch_lm(m, K, o.lm_fp64, KIND_MULTI, f0, c->stream));
    }
    if (o.rician) {
      // tau from the LM fits (recomputed for every voxel; only the rerun ones read it again)
      TauArgs t{};
      t.data = c->d_data.p;
      t.N = c->N;
      t.nvox = nv;
      t.acq = c->d_acq.p;
      t.params = c->d_params.p;
      t.np = np;
      t.model = o.model;
      t.f0 = f0;
      t.R = o.R_prior_mean;
      t.tau = c->d_tau.p;
      HIPCHK(c, launch_tau(t, K, o.lm_fp64, c->stream));
    }
    // their chains with the configured burn-in and default widths
    McmcArgs b = a;
    b.nburn = o.nburn;
    b.widths_in = nullptr;
    b.list = c->d_list.p;
    b.nvox = nfb;
    HIPCHK(c, fn(&b, K, o.rician, c->stream));
    if (int rc = finish_async(c)) return rc;
  }
  HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  c->timed_mcmc = true;
  c->timed_finish = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dir.release();
  c->use_widths_in = false;
  c->have_finish = true;
  return FDT_OK;
}

int fdt_get_fallback_flags(fdt_ctx* c, unsigned char* flags) {
  if (!c || !flags) return FDT_EINVAL;
  if (!c->d_refit.p || c->d_refit.n < (size_t)c->nvox) return set_err(c, FDT_ESTATE, "no learned-start run");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipMemcpyAsync(flags, c->d_refit.p, c->nvox, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_set_widths(fdt_ctx* c, const float* widths) {
  if (!c) return FDT_EINVAL;
  if (!widths) { c->use_widths_in = false; return FDT_OK; }
  (void)hipSetDevice(c->device);
  HIPCHK(c, c->d_widths_in.ensure((size_t)std::max(c->nvox, 1) * 16));
  if (c->nvox)
    HIPCHK(c, hipMemcpyAsync(c->d_widths_in.p, widths, sizeof(float) * c->nvox * 16, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->use_widths_in = true;
  return FDT_OK;
}

int fdt_get_energy(fdt_ctx* c, float* energy) {
  if (!c || !energy) return FDT_EINVAL;
  if (!c->have_samples || !c->energy_valid) return set_err(c, FDT_ESTATE, "no chains have run");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipMemcpyAsync(energy, c->d_energy.p, sizeof(float) * c->nvox, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

int fdt_get_widths(fdt_ctx* c, float* widths) {
  if (!c || !widths) return FDT_EINVAL;
  if (!c->have_samples) return set_err(c, FDT_ESTATE, "no chains have run");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipMemcpyAsync(widths, c->d_widths_out.p, sizeof(float) * c->nvox * 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FDT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ prediction
// xfibres_pred (xfibres_pred.cc:153-190) on the device: upload the posterior
// means, one launch, download the [N][nvox] prediction.
extern "C" {

int fdt_predict(int device, int N, const float* bvals, const float* bvecs, int nvox, int K, int model, const float* S0,
                const float* d, const float* dstd, const float* R, const float* f0, const float* f, const float* dyads,
                float* out) {
  if (N < 1 || nvox < 0 || K < 0 || model < 1 || model > 3) return FDT_EINVAL;
  if (!bvals || !bvecs || !out || (nvox > 0 && (!S0 || !d || (K > 0 && (!f || !dyads)))) ||
      (model >= 2 && nvox > 0 && !dstd) || (model == 3 && nvox > 0 && !R))
    return FDT_EINVAL;
  if (nvox == 0) return FDT_OK;
  if (hipSetDevice(device) != hipSuccess) return FDT_EHIP;
  const size_t nv = (size_t)nvox;
  DevBuf<float> bv, bc, s0, dd, ds, rr, ff0, ff, dy, o;
  hipStream_t st = nullptr;
  auto up = [&](DevBuf<float>& b, const float* h, size_t n) -> bool {
    if (!h) return true;
    return b.ensure(n) == hipSuccess && hipMemcpyAsync(b.p, h, n * sizeof(float), hipMemcpyHostToDevice, st) == hipSuccess;
  };
  bool ok = hipStreamCreate(&st) == hipSuccess;
  ok = ok && up(bv, bvals, N) && up(bc, bvecs, 3 * (size_t)N) && up(s0, S0, nv) && up(dd, d, nv);
  ok = ok && up(ds, model >= 2 ? dstd : nullptr, nv) && up(rr, model == 3 ? R : nullptr, nv) && up(ff0, f0, nv);
  ok = ok && up(ff, K ? f : nullptr, (size_t)K * nv) && up(dy, K ? dyads : nullptr, (size_t)K * 3 * nv);
  ok = ok && o.ensure((size_t)N * nv) == hipSuccess;
  if (ok) {
    PredArgs a{};
    a.N = N; a.nvox = nvox; a.K = K; a.model = model;
    a.bvals = bv.p; a.bvecs = bc.p; a.S0 = s0.p; a.d = dd.p; a.dstd = ds.p; a.R = rr.p; a.f0v = f0 ? ff0.p : nullptr;
    a.f = ff.p; a.dyads = dy.p; a.out = o.p;
    ok = fdt_launch_pred(&a, st) == hipSuccess &&
         hipMemcpyAsync(out, o.p, (size_t)N * nv * sizeof(float), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
  }
  if (st) (void)hipStreamDestroy(st);
  for (auto* b : {&bv, &bc, &s0, &dd, &ds, &rr, &ff0, &ff, &dy, &o}) b->release();
  return ok ? FDT_OK : FDT_EHIP;
}

}  // extern "C"
