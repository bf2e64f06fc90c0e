// fdt_api.hip -- host orchestration behind the C ABI of include/fdt_xfibres.h.
//
// Replaces the reference's per-sub-part host pipeline (xfibres_gpu(),
// CUDA/xfibres_gpu.cu:29-146) with a context that owns one stream and the
// device buffers; data, LM parameters and samples stay resident between
// calls.  No call exits the process (the reference's sync_check does,
// CUDA/sync_check.h:18-25); errors come back as FDT_E* codes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fdt_xfibres.h"
#include "fdt_kernels.h"

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
hipError_t fdt_launch_f0_flag(const F0FlagArgs*, hipStream_t);
hipError_t fdt_launch_zero_ints(int*, int, hipStream_t);
hipError_t fdt_launch_mcmc_queue(const QueueArgs*, int rows_per_wg, int max_items, hipStream_t);
hipError_t fdt_launch_finish(const FinishArgs*, hipStream_t);
hipError_t fdt_launch_transpose(const float*, float*, int, int, hipStream_t);
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
  DevBuf<unsigned char> d_refit;  // voxels refitted without f0
  bool have_params = false, have_samples = false, have_finish = false;
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
  return FDT_OK;
}

extern "C" {

void fdt_default_opts(fdt_opts* o) {
  std::memset(o, 0, sizeof *o);
  o->nfibres = 1;
  o->model = 1;
  o->fudge = 1.0f;
  o->njumps = 1250;
  o->nburn = 1000;
  o->sampleevery = 25;
  o->updateproposalevery = 40;
  o->seed = 8665904;
  o->R_prior_mean = 0.13f;
  o->R_prior_std = 0.03f;
  o->R_prior_fudge = 0.0f;
}

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
  c->d_shell.release();
  c->d_shell_b.release();
  c->d_iters.release();
  c->d_queue.release();
  c->d_list.release();
  c->d_s1.release();
  c->d_refit.release();
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
  // b-shells: distinct b values (exact), used by the MCMC when there are at most 16
  std::vector<float> shb;
  std::vector<int> shell(c->Npad, 0);
  for (int i = 0; i < N && shb.size() <= 16; ++i) {
    auto it = std::find(shb.begin(), shb.end(), bvals[i]);
    if (it == shb.end()) { shb.push_back(bvals[i]); shell[i] = (int)shb.size() - 1; }
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
  c->have_params = c->have_samples = c->have_finish = false;
  return FDT_OK;
}

static hipError_t launch_lm(const LmArgs& a, int K, int fp64, int kind, int f0, hipStream_t s) {
  if (!fp64) {
    if (K == 1) return fdt_launch_lm_k1_0(&a, kind, f0, s);
    if (K == 2) return fdt_launch_lm_k2_0(&a, kind, f0, s);
    if (K == 3) return fdt_launch_lm_k3_0(&a, kind, f0, s);
  } else {
    if (K == 1) return fdt_launch_lm_k1_1(&a, kind, f0, s);
    if (K == 2) return fdt_launch_lm_k2_1(&a, kind, f0, s);
    if (K == 3) return fdt_launch_lm_k3_1(&a, kind, f0, s);
  }
  return hipErrorInvalidValue;
}

static hipError_t launch_tau(const TauArgs& a, int K, int fp64, hipStream_t s) {
  if (!fp64) {
    if (K == 1) return fdt_launch_tau_k1_0(&a, s);
    if (K == 2) return fdt_launch_tau_k2_0(&a, s);
    if (K == 3) return fdt_launch_tau_k3_0(&a, s);
  } else {
    if (K == 1) return fdt_launch_tau_k1_1(&a, s);
    if (K == 2) return fdt_launch_tau_k2_1(&a, s);
    if (K == 3) return fdt_launch_tau_k3_1(&a, s);
  }
  return hipErrorInvalidValue;
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
  return a;
}

// The fit() pipeline of CUDA/xfibres_gpu.cu:173-348 as a chain of launches
// on the context stream: stage 1 (PVM_single, or PVM_single_c for model 1
// without --nonlinear), PVM_multi for model >= 2, the --f0 refit of
// implausible voxels over a device list (model 1: the same stage without f0;
// model >= 2: PVM_single_c then PVM_multi, :331-333), and tau for --rician.
static int fit_async(fdt_ctx* c) {
  if (c->N <= 0) return set_err(c, FDT_ESTATE, "no acquisition");
  const fdt_opts& o = c->o;
  const int K = o.nfibres, fp64 = o.lm_fp64 ? 1 : 0, f0 = o.f0 ? 1 : 0;
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
    HIPCHK(c, launch_tau(t, K, fp64, c->stream));
  }
  if (!c->quiet) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  c->timed_fit = !c->quiet;
  c->have_params = true;
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
  HIPCHK(c, launch_lm(a, K, c->o.lm_fp64 ? 1 : 0, kind, f0 ? 1 : 0, c->stream));
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
  HIPCHK(c, c->d_queue.ensure(NQUEUE));
  HIPCHK(c, fdt_launch_zero_ints(c->d_queue.p, NQUEUE, c->stream));
  LmArgs a = base_lm_args(c);
  a.init = INIT_DTI;
  a.lm = 0;
  a.queue = c->d_queue.p;
  a.out = dout.p;
  a.out_np = np;
  HIPCHK(c, launch_lm(a, K, c->o.lm_fp64 ? 1 : 0, kind, f0 ? 1 : 0, c->stream));
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
  c->have_params = true;
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
  int rc = fit_async(c);
  if (!rc) rc = mcmc_async(c, vox_offset);
  if (!rc) rc = finish_async(c);
  return rc;
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

static int tile_of(int N, int force) {
  for (const auto& e : kVpl)
    if (16 * e.vpl >= N && e.vpl >= force) return e.vpl;
  return 0;
}

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

static bool in_queue_kernel(const fdt_ctx* c) {
  int vpl = tile_of(c->N, c->vpl_force);
  return !c->o.rician && vpl > 0 && vpl <= QUEUE_MAX_VPL;
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
