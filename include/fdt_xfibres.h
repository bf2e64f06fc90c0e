/* fdt_xfibres.h -- C ABI of the MI355X xfibres engine (libfdt_xfibres.so).
 *
 * The engine replaces the reference's GPU library libbedpostx_cuda.so behind
 * the xfibres_gpu host driver.  The reference library exports one C++ entry,
 *     void xfibres_gpu(Matrix datam, Matrix bvecs, Matrix bvals, Matrix gradm,
 *                      int idpart, int idSubpart, float seed, string subjdir)
 * (CUDA/xfibres_gpu.h:15-23, CUDA/xfibres_gpu.cu:29-146), reads its options
 * from the xfibresOptions singleton (CUDA/xfibres_gpu.cu:67) and returns its
 * results only as files (CUDA/samples.cu:231-347).  Here the same pipeline is
 * split into plain-pointer calls so a caller controls residency:
 *
 *   fdt_create / fdt_set_acquisition / fdt_load_data   -- the inputs
 *   fdt_fit      -- LM initialisation  (CUDA/xfibres_gpu.cu:173-348 fit(),
 *                   CUDA/diffmodels.cu:26-347, the three PVM LM kernels)
 *   fdt_mcmc     -- burn-in + record   (CUDA/runmcmc.cu:27-316,
 *                   CUDA/runmcmc_kernels.cu:193-1474)
 *   fdt_finish   -- fibre sort, means, dyads (CUDA/samples.cu:134-347,
 *                   make_dyadic_vectors.cc:60-88)
 *   fdt_get_outputs -- device -> caller buffers (CUDA/xfibres_gpu.cu:659-738)
 *
 * Conventions: every call returns 0 or a negative FDT_E* code and never exits;
 * fdt_last_error() gives the message.  One context per host thread; a context
 * owns one HIP stream and all device buffers; caller buffers are host memory
 * and are only read or written during the call that takes them.
 *
 * Sample layout (host and device): per-fibre samples q[k][s][v]
 * (K x nsamples x nvox); per-voxel scalars q[s][v].  Each q[k] block is the
 * reference's [nsamples][nvox] row-major "<name>samples" part file.
 */
#ifndef FDT_XFIBRES_H
#define FDT_XFIBRES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDT_OK 0
#define FDT_EINVAL (-1)   /* bad argument or shape */
#define FDT_EHIP (-2)     /* HIP runtime error */
#define FDT_ESTATE (-3)   /* call out of order (e.g. mcmc before fit) */
#define FDT_ENOMEM (-4)   /* device allocation failed */
#define FDT_EUNSUP (-5)   /* option combination not built */

#define FDT_LAYOUT_VOX_MAJOR 0 /* data[v*N + i] */
#define FDT_LAYOUT_VOL_MAJOR 1 /* data[i*nvox + v]: the reference data_<part> file layout */

/* The xfibres option surface (xfibresoptions.h:25-56, defaults :104-148). */
typedef struct fdt_opts {
  int nfibres;             /* --nf (1..3) */
  int model;               /* --model 1/2/3 */
  float fudge;             /* --fudge */
  int njumps;              /* --nj */
  int nburn;               /* --bi */
  int sampleevery;         /* --se */
  int updateproposalevery; /* --upe */
  uint64_t seed;           /* --seed */
  int no_ard;              /* --noard */
  int all_ard;             /* --allard */
  int nonlin;              /* --nonlinear */
  int cnonlin;             /* --cnonlinear */
  int rician;              /* --rician */
  int f0;                  /* --f0 */
  int ardf0;               /* --ardf0 */
  float R_prior_mean;      /* --Rmean */
  float R_prior_std;       /* --Rstd */
  float R_prior_fudge;     /* --Rfudge */
  int lm_fp64;             /* LM arithmetic: 0 FP32 model, Jacobian and cost; 1 FP64 throughout;
                              2 (default) FP32 Jacobian and normal equations with FP64 parameters and cost
                              (every accept/reject and stop decision on an FP64 cost); with --f0, 2 runs
                              as 1 */
  int cpu_init;            /* 1: start the chains as the CPU xfibres does (xfibres.cc:517-541): without
                              nonlinear fitting for Gaussian noise when neither nonlin nor cnonlin is set,
                              from the tensor (nospat) or from the best finished neighbour (needs the
                              voxel grid, fdt_set_grid); 0: the xfibres_gpu fit (CUDA/xfibres_gpu.cu:173-348) */
  int nospat;              /* --nospat (xfibresoptions.h:134): tensor start, no neighbour initialisation */
  int spatial_serial;      /* neighbour initialisation one voxel per step in mask order, the CPU's own order
                              (checks the wavefront schedule, which gives identical results); 0: wavefronts */
  int reserved[4];
} fdt_opts;

/* Output buffers; any pointer may be NULL to skip that output. */
typedef struct fdt_outputs {
  float* th;        /* K x ns x nvox, fibres sorted by descending sum_s f */
  float* ph;        /* K x ns x nvox */
  float* f;         /* K x ns x nvox */
  float* mean_S0;   /* nvox */
  float* mean_d;    /* nvox */
  float* mean_dstd; /* nvox (model >= 2) */
  float* mean_R;    /* nvox (model 3) */
  float* mean_f0;   /* nvox (--f0) */
  float* mean_tau;  /* nvox (--rician) */
  float* mean_f;    /* K x nvox */
  float* dyads;     /* K x 3 x nvox: principal eigenvector of the mean dyadic tensor */
  float* disp;      /* K x nvox: 1 - largest eigenvalue of the mean dyadic tensor */
  float* S0;        /* ns x nvox raw samples (optional) */
  float* d;         /* ns x nvox */
  float* dstd;      /* ns x nvox */
  float* R;         /* ns x nvox */
  float* f0;        /* ns x nvox */
  float* tau;       /* ns x nvox */
} fdt_outputs;

typedef struct fdt_timing {
  float ms_fit;     /* device time of the LM kernels of the last fdt_fit */
  float ms_mcmc;    /* device time of the MCMC kernel of the last fdt_mcmc */
  float ms_finish;  /* device time of the finish kernel of the last fdt_finish */
  int mcmc_vpl;     /* volumes per lane chosen for the MCMC kernel */
  int mcmc_lanes;   /* lanes per voxel */
} fdt_timing;

typedef struct fdt_ctx fdt_ctx;

/* Fill o with the xfibres defaults (xfibresoptions.h:104-148). */
void fdt_default_opts(fdt_opts* o);

/* Create a context on HIP device `device` (replaces init_gpu,
 * CUDA/init_gpu.cu:12-25, and the options singleton). NULL on failure. */
fdt_ctx* fdt_create(int device, const fdt_opts* o);
void fdt_destroy(fdt_ctx* c);
const char* fdt_last_error(const fdt_ctx* c);

/* bvals [N], bvecs 3 x N row-major (x row, y row, z row), already unit
 * length as xfibres_gpu.cc:41-53 leaves them (replaces the bvals/bvecs
 * arguments of xfibres_gpu(), CUDA/xfibres_gpu.h:17-18). */
int fdt_set_acquisition(fdt_ctx* c, int N, const float* bvals, const float* bvecs);

/* Copy nvox voxels of N volumes to the device (replaces datam,
 * CUDA/xfibres_gpu.h:16, and prepare_data_gpu_FIT, CUDA/xfibres_gpu.cu:370-461).
 * `layout` is FDT_LAYOUT_VOL_MAJOR for the data_<part> file order. */
int fdt_load_data(fdt_ctx* c, const float* data, int layout, int nvox);

/* --gradnonlin: the gradient-deviation tensor of every voxel, 9 values
 * (L(i,j) = g[3j+i]) per voxel; FDT_LAYOUT_VOL_MAJOR is the [9][nvox]
 * grad_dev_<part> file order (replaces gradm, CUDA/xfibres_gpu.h:19, read by
 * xfibres_gpu.cc:115-122).  Every later fit and chain then uses each voxel's
 * corrected bvecs/bvals (correct_bvals_bvecs, CUDA/xfibres_gpu.cu:151-170,
 * per-voxel acquisitions :403-436), built on the device.  nvox must equal the
 * loaded data's.  grad == NULL returns to the shared acquisition. */
int fdt_load_grad_dev(fdt_ctx* c, const float* grad, int layout, int nvox);

/* The acquisition voxel `vox` is fitted with: bvals [N], bvecs 3 x N (the
 * shared one, or its corrected copy under --gradnonlin).  No counterpart in
 * the reference; for inspection and tests. */
int fdt_get_voxel_acquisition(fdt_ctx* c, int vox, float* bvals, float* bvecs);

/* Number of LM output parameters per voxel: 2 + 3K (+1 model>=2) (+1 --f0);
 * layout [S0, d, (d_std), f_k, th_k, ph_k ..., (f0)] (CUDA/xfibres_gpu.cu:450-454). */
int fdt_num_params(const fdt_ctx* c);

/* LM initialisation on the resident data (CUDA/xfibres_gpu.cu:173-348 and
 * calculate_tau, CUDA/diffmodels.cu:259-347).  Tensor start values are
 * computed on the device.  Results stay resident. */
int fdt_fit(fdt_ctx* c);

/* One LM stage from caller-given start values, for parity checks of the LM
 * kernels alone (fit_PVM_single_kernel CUDA/PVM_single.cu:387-473,
 * fit_PVM_single_c_kernel CUDA/PVM_single_c.cu:516-616,
 * fit_PVM_multi_kernel CUDA/PVM_multi.cu:468-576).
 * kind 0 PVM_single, 1 PVM_single_c, 2 PVM_multi (G = Gamma_for_ball_only,
 * R = model-3 ratio); start/out [nvox][np] with np = (kind==2?3:2)+3K+f0;
 * start in the LM's transformed space, out finalised (as the kernels leave
 * it). iters may be NULL. Uses the resident data. */
int fdt_lm_stage(fdt_ctx* c, int kind, int K, int f0, int G, float R, const float* start, float* out, int* iters);

/* Device tensor start values of one LM stage (CUDA/diffmodels.cu:57-93,
 * 149-183): out [nvox][np] for kind 0 (PVM_single) or 1 (PVM_single_c). */
int fdt_lm_start(fdt_ctx* c, int kind, int K, int f0, float* out);

/* Read / replace the LM result that starts the chains: params [nvox][np]
 * and tau [nvox] (may be NULL).  In the reference this hand-over is internal
 * to xfibres_gpu() (fit() output into init_Fibres_Multifibres_kernel,
 * CUDA/xfibres_gpu.cu:88-128); it is exposed so a caller can start the chains
 * from its own initialisation. */
int fdt_get_params(fdt_ctx* c, float* params, float* tau);

/* LM loop trips per voxel (successful and failed iterations, summed over the
 * stages and the --f0 refit) of the last fdt_fit / fdt_run: iters [nvox].  The
 * reference counts them inside levenberg_marquardt_PVM_*_gpu
 * (CUDA/levenberg_marquardt.cu:24-333) without reporting them; they are
 * exposed for the run report (iteration histogram).  FDT_ESTATE when the
 * current parameters were not produced by the LM fit. */
int fdt_get_lm_iters(fdt_ctx* c, int* iters);
int fdt_set_params(fdt_ctx* c, const float* params, const float* tau);

/* Burn-in + record as one persistent launch (CUDA/runmcmc.cu:108-316).
 * Draws are Philox4x32-10 keyed by --seed and countered by
 * (vox_offset + voxel, jump, proposal), so results do not depend on how the
 * voxels are split into calls. */
int fdt_mcmc(fdt_ctx* c, uint64_t vox_offset);

/* Sort fibres by the sum of their f samples, means, dyads and dispersion on
 * the device (record_finish_voxels, CUDA/xfibres_gpu.cu:659-738, and
 * Samples::record/finish_voxel/save, CUDA/samples.cu:20-347). */
int fdt_finish(fdt_ctx* c);
/* Copy the finished outputs into caller-owned host buffers (the arrays that
 * Samples::save writes as <name>J files, CUDA/samples.cu:262-347); NULL
 * fields are skipped. */
int fdt_get_outputs(fdt_ctx* c, const fdt_outputs* out);

/* fit + mcmc + finish on the resident data, asynchronous on the context
 * stream: the per-part pipeline of xfibres_gpu(), CUDA/xfibres_gpu.cu:29-146.
 * With cpu_init and neighbour initialisation it is the voxel loop of the CPU
 * xfibres instead (xfibres.cc:898-921): each chain starts from the means of
 * its best finished neighbour (Samples::neighbour_initialise, xfibres.cc:306-357),
 * run on the device in wavefronts x + 2y + 4z, which order every voxel after
 * the neighbours that precede it in the mask; fdt_fit is then refused. */
int fdt_run(fdt_ctx* c, uint64_t vox_offset);

/* The grid of the loaded voxels for neighbour initialisation: nx, ny, nz and
 * each voxel's index x + nx (y + ny z), strictly increasing (mask order, the
 * columns of volume4D::matrix(mask) that xfibres.cc:874-876 reads with
 * vol2matrixkey / matrix2volkey).  No counterpart in the xfibres_gpu boundary. */
int fdt_set_grid(fdt_ctx* c, int nx, int ny, int nz, const long long* gidx);

/* After a neighbour-initialised fdt_run: what each finished voxel left for its
 * neighbours (Samples::finish_voxel, xfibres.cc:249-303), in chain fibre order:
 * been [nvox] fibres with mean f > 0.05, mean_f [K][nvox], mean_d and mean_S0
 * [nvox], dyad [K][3][nvox] (principal eigenvector of the summed dyadic tensor,
 * largest component positive).  NULL fields are skipped. */
int fdt_get_spatial_state(fdt_ctx* c, int* been, float* mean_f, float* mean_d, float* mean_S0, double* dyad);
int fdt_sync(fdt_ctx* c);
int fdt_get_timing(fdt_ctx* c, fdt_timing* t);

/* Force the MCMC volumes-per-lane tile (0 = automatic).  No counterpart in
 * the reference (its MCMC block shape is fixed, CUDA/runmcmc.cu:108-316). */
int fdt_set_mcmc_tile(fdt_ctx* c, int vpl);

/* ---------------------------------------------------------------- batches
 * A queue of volumes with their own acquisitions, models and fibre counts,
 * run together (the reference runs one volume per process and sub-part,
 * CUDA/xfibres_gpu.cu:29-146, and has no batching): the LM stages of every
 * volume run concurrently, then one persistent chain launch covers every
 * volume whose tile the queue kernel carries (Gaussian noise, N <= 288),
 * then the finishing passes.  The first fdt_run_queue runs eagerly and
 * records the launches into a hipGraph that later runs replay. */
typedef struct fdt_queue fdt_queue;
fdt_queue* fdt_queue_create(int device);
void fdt_queue_destroy(fdt_queue* q);
const char* fdt_queue_last_error(const fdt_queue* q);
/* Add a volume: its options, acquisition and data (copied to the device).
 * Returns the job index (>= 0) or a negative FDT_E* code. */
int fdt_enqueue_batch(fdt_queue* q, const fdt_opts* o, int N, const float* bvals, const float* bvecs,
                      const float* data, int layout, int nvox, uint64_t vox_offset);
int fdt_run_queue(fdt_queue* q);
int fdt_queue_num_jobs(const fdt_queue* q);
int fdt_queue_get_outputs(fdt_queue* q, int job, const fdt_outputs* out);
int fdt_queue_get_params(fdt_queue* q, int job, float* params, float* tau);
/* --gradnonlin for job `job` of the queue: its gradient deviations
 * ([9][nvox] with FDT_LAYOUT_VOL_MAJOR, as fdt_load_grad_dev; NULL clears).
 * Such a job runs its chains in its own launch inside the batch.  No
 * counterpart in the reference (its parts run one process each). */
int fdt_queue_set_grad_dev(fdt_queue* q, int job, const float* grad, int layout);

/* Voxels per workgroup of the queue kernel (4, 8, 16; 0 = keep) and the
 * volumes-per-lane tile of one job (job < 0: none; vpl 0 = automatic). */
int fdt_queue_set_tiles(fdt_queue* q, int rows_per_wg, int job, int vpl);
/* Time the queue kernel on the first sample_vox voxels of every volume for
 * `jumps` sweeps over the tile choices and keep the fastest. */
int fdt_queue_autotune(fdt_queue* q, int jumps, int sample_vox, float* ms_best);
/* Duration of the last fdt_run_queue (ms) and the tiles in use. */
int fdt_queue_timing(fdt_queue* q, float* ms_last, int* rows_per_wg, int* vpl_of_job, int njobs);

/* Final proposal width of every voxel and slot after fdt_mcmc / fdt_run:
 * widths [nvox][16] (slots f0, tau, d, d_std, R, S0, th_k, ph_k, f_k).  No
 * counterpart in the reference (its widths stay in FibreGPU,
 * CUDA/fibre_gpu.h:12-83); they label the training set of the learned start. */
int fdt_get_widths(fdt_ctx* c, float* widths);

/* Initial proposal widths for the next chains: widths [nvox][16] per voxel and
 * slot (layout of fdt_get_widths); a value <= 0 keeps the default width of
 * init_Fibres_Multifibres_kernel (CUDA/runmcmc_kernels.cu:497-503).  NULL
 * clears them.  Cleared by fdt_load_data; the learned start
 * (fdt_run_surrogate) sets its own. */
int fdt_set_widths(fdt_ctx* c, const float* widths);

/* Per voxel, the likelihood energy each chain carries at its last sweep of
 * the last fdt_mcmc / fdt_run: Gaussian noise: the running sum of squared
 * residuals; Rician: -N log tau minus the summed log-likelihood terms (the
 * likelihood part of compute_likelihood, CUDA/runmcmc_kernels.cu:147-191).
 * The chains update it incrementally; exposed to bound that drift. */
int fdt_get_energy(fdt_ctx* c, float* energy);

/* ------------------------------------------------------ learned start
 * An MLP N -> 512 -> 512 -> 256 -> 28 (GELU, BF16 weights and activations,
 * FP32 accumulation) maps a voxel's signal over its mean b=0 signal to start
 * values (3 f, 3 directions, d, d_std) and 14 log proposal widths.  It is
 * trained on the GPU on labels from this engine's own LM + MCMC.  No
 * counterpart in the reference. */
typedef struct fdt_surrogate fdt_surrogate;
fdt_surrogate* fdt_surrogate_create(int device, int n_in, uint64_t seed);
void fdt_surrogate_destroy(fdt_surrogate* s);
const char* fdt_surrogate_last_error(const fdt_surrogate* s);
size_t fdt_surrogate_num_weights(const fdt_surrogate* s);
/* flat FP32 weights: W1 [N][512], b1, W2 [512][512], b2, W3 [512][256], b3, W4 [256][28], b4 */
int fdt_surrogate_load(fdt_surrogate* s, const float* w, size_t n);
int fdt_surrogate_save(fdt_surrogate* s, float* w, size_t n);
/* training set: raw signals [n][N] (voxel-major), bvals [N], targets [n][28]:
 * f_1..3, dyads x_1..3 (3 each), 1000 d, 1000 d_std, log widths of slots
 * f0, d, d_std, R, S0, th1, ph1, f1, th2, ph2, f2, th3, ph3, f3 (>= 1e29: unused) */
int fdt_surrogate_set_training_data(fdt_surrogate* s, const float* signals, const float* bvals, const float* targets,
                                    int n);
/* gradients (flat, as the weights) of the mean loss over rows [first, first+batch) */
int fdt_surrogate_gradients(fdt_surrogate* s, int first, int batch, float* grads, float* loss);
/* one Adam step (lr; beta 0.9/0.999) on rows [first, first+batch) */
int fdt_surrogate_train_step(fdt_surrogate* s, int first, int batch, float lr, float* loss);
/* outputs [n][28] for raw signals [n][N] */
int fdt_surrogate_forward(fdt_surrogate* s, const float* signals, const float* bvals, int n, float* out);
/* start values and proposal widths of the context's resident voxels from the
 * network (replaces fdt_fit) */
int fdt_surrogate_infer(fdt_surrogate* s, fdt_ctx* c);
/* learned start + nburn_short burn-in, fallback of suspicious voxels to the
 * plain LM + configured burn-in, finishing; *n_fallback = rerun voxels */
int fdt_run_surrogate(fdt_ctx* c, fdt_surrogate* s, uint64_t vox_offset, int nburn_short, float chi2_max,
                      float ang_max_deg, int* n_fallback);
/* why each voxel of the last fdt_run_surrogate was rerun: [nvox] bit 1 sum of
 * squares, bit 2 direction (0: kept) */
int fdt_get_fallback_flags(fdt_ctx* c, unsigned char* flags);

/* ------------------------------------------------------ posterior prediction
 * The signal predicted by the posterior means (xfibres_pred.cc:153-190, the
 * per-voxel loop of xfibres_pred): for every volume l and voxel i,
 * S0 (sum_k f_k stick_k + (1 - sum f - f0) ball + f0) with the model's stick
 * and ball forms (model 1; model 2 gamma diffusivities; model 3 zeppelins),
 * evaluated on device `device`.  Inputs per voxel [nvox] (dstd for model >= 2,
 * R for model 3, f0 optional), f [K][nvox], dyads [K][3][nvox]; bvals [N],
 * bvecs 3 x N unit vectors; out [N][nvox]. */
int fdt_predict(int device, int N, const float* bvals, const float* bvecs, int nvox, int K, int model, const float* S0,
                const float* d, const float* dstd, const float* R, const float* f0, const float* f, const float* dyads,
                float* out);

/* Allocation-free chains and finishing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Runs what
 * fdt_mcmc + fdt_finish run, for nvox voxels held in the caller's device
 * memory, without a context: d_bvals [N] and d_bvecs [3][N] (as
 * fdt_set_acquisition takes them), d_data [nvox][N] voxel-major, d_params
 * [nvox][np] (the layout of fdt_get_params), d_tau [nvox] (--rician, else
 * NULL).  Every pointer of *out is a device pointer sized as for
 * fdt_get_outputs; th/ph/f, the means, dyads, disp and the scalar samples of
 * the model's parameters are required (the scalar samples feed the means).
 * The library's scratch -- the acquisition table and the unsorted fibre
 * samples -- is a static device array on each GPU (held until the process
 * ends): 3 * nfibres * (njumps / sampleevery) * nvox must not exceed
 * FDT_NOALLOC_SAMPLE_FLOATS, else FDT_EINVAL.  Calls no hipMalloc/hipFree;
 * every launch is ordered on `stream` (a hipStream_t; NULL is the null
 * stream) and nothing synchronises the host.  Calls that overlap on one GPU
 * share the scratch array and race.  On error the message is written to msg
 * (msg_len bytes, may be NULL).  Outputs equal those of fdt_set_params +
 * fdt_mcmc + fdt_finish + fdt_get_outputs on the same inputs. */
#define FDT_NOALLOC_SAMPLE_FLOATS (4u << 20) /* 16 MiB of unsorted fibre samples */
int fdt_mcmc_noalloc(const fdt_opts* o, int N, const float* d_bvals, const float* d_bvecs, const float* d_data,
                     const float* d_params, const float* d_tau, uint64_t vox_offset, int nvox,
                     const fdt_outputs* out, void* stream, char* msg, size_t msg_len);

#ifdef __cplusplus
}
#endif
#endif
