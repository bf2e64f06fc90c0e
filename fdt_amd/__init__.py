"""fdt_amd -- MI355X xfibres engine (Bayesian ball-and-sticks estimation).

Python view of the C ABI in include/fdt_xfibres.h (libfdt_xfibres.so).  The
engine is the HIP library; this module only moves numpy arrays across the
boundary.  Loading fails loudly when the library has not been built: there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdt_xfibres.so")
CLI_PATH = os.path.join(HERE, "bin", "xfibres_gpu")

LAYOUT_VOX_MAJOR, LAYOUT_VOL_MAJOR = 0, 1
KIND_SINGLE, KIND_SINGLE_C, KIND_MULTI = 0, 1, 2

EXPORTS = ["fdt_default_opts", "fdt_create", "fdt_destroy", "fdt_last_error", "fdt_set_acquisition",
           "fdt_load_data", "fdt_load_grad_dev", "fdt_get_voxel_acquisition", "fdt_num_params", "fdt_fit", "fdt_lm_stage", "fdt_lm_start", "fdt_get_params",
           "fdt_set_params", "fdt_mcmc", "fdt_finish", "fdt_get_outputs", "fdt_run", "fdt_sync",
           "fdt_get_timing", "fdt_set_mcmc_tile",
           "fdt_queue_create", "fdt_queue_destroy", "fdt_queue_last_error", "fdt_enqueue_batch", "fdt_run_queue",
           "fdt_queue_num_jobs", "fdt_queue_get_outputs", "fdt_queue_get_params", "fdt_queue_set_tiles",
           "fdt_queue_autotune", "fdt_queue_timing", "fdt_get_widths", "fdt_queue_set_grad_dev",
           "fdt_surrogate_create", "fdt_surrogate_destroy", "fdt_surrogate_last_error", "fdt_surrogate_num_weights",
           "fdt_surrogate_load", "fdt_surrogate_save", "fdt_surrogate_set_training_data", "fdt_surrogate_gradients",
           "fdt_surrogate_train_step", "fdt_surrogate_forward", "fdt_surrogate_infer", "fdt_run_surrogate", "fdt_get_fallback_flags",
           "fdt_predict", "fdt_set_grid", "fdt_get_spatial_state", "fdt_get_lm_iters", "fdt_mcmc_noalloc", "fdt_set_widths", "fdt_get_energy"]


class Opts(C.Structure):
    _fields_ = [
        ("nfibres", C.c_int), ("model", C.c_int), ("fudge", C.c_float), ("njumps", C.c_int), ("nburn", C.c_int),
        ("sampleevery", C.c_int), ("updateproposalevery", C.c_int), ("seed", C.c_uint64), ("no_ard", C.c_int),
        ("all_ard", C.c_int), ("nonlin", C.c_int), ("cnonlin", C.c_int), ("rician", C.c_int), ("f0", C.c_int),
        ("ardf0", C.c_int), ("R_prior_mean", C.c_float), ("R_prior_std", C.c_float), ("R_prior_fudge", C.c_float),
        ("lm_fp64", C.c_int), ("cpu_init", C.c_int), ("nospat", C.c_int), ("spatial_serial", C.c_int),
        ("reserved", C.c_int * 4),
    ]


_FP = C.POINTER(C.c_float)
_OUT_FIELDS = ["th", "ph", "f", "mean_S0", "mean_d", "mean_dstd", "mean_R", "mean_f0", "mean_tau", "mean_f", "dyads",
               "disp", "S0", "d", "dstd", "R", "f0", "tau"]


class Outputs(C.Structure):
    _fields_ = [(n, _FP) for n in _OUT_FIELDS]


class Timing(C.Structure):
    _fields_ = [("ms_fit", C.c_float), ("ms_mcmc", C.c_float), ("ms_finish", C.c_float), ("mcmc_vpl", C.c_int),
                ("mcmc_lanes", C.c_int)]


_lib = None


def build(jobs: int = 8) -> None:
    """Compile the HIP library and the xfibres_gpu driver for gfx950 (in-tree)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run fdt_amd.build() (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        vp, ip = C.c_void_p, C.POINTER(C.c_int)
        L.fdt_default_opts.argtypes = [C.POINTER(Opts)]
        L.fdt_default_opts.restype = None
        L.fdt_create.argtypes = [C.c_int, C.POINTER(Opts)]
        L.fdt_create.restype = vp
        L.fdt_destroy.argtypes = [vp]
        L.fdt_destroy.restype = None
        L.fdt_last_error.argtypes = [vp]
        L.fdt_last_error.restype = C.c_char_p
        L.fdt_set_acquisition.argtypes = [vp, C.c_int, _FP, _FP]
        L.fdt_load_data.argtypes = [vp, _FP, C.c_int, C.c_int]
        L.fdt_load_grad_dev.argtypes = [vp, _FP, C.c_int, C.c_int]
        L.fdt_get_voxel_acquisition.argtypes = [vp, C.c_int, _FP, _FP]
        L.fdt_predict.argtypes = [C.c_int, C.c_int, _FP, _FP, C.c_int, C.c_int, C.c_int] + [_FP] * 8
        L.fdt_num_params.argtypes = [vp]
        L.fdt_fit.argtypes = [vp]
        L.fdt_lm_stage.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, _FP, _FP, ip]
        L.fdt_lm_start.argtypes = [vp, C.c_int, C.c_int, C.c_int, _FP]
        L.fdt_get_params.argtypes = [vp, _FP, _FP]
        L.fdt_get_lm_iters.argtypes = [vp, ip]
        L.fdt_mcmc_noalloc.argtypes = [C.POINTER(Opts), C.c_int, vp, vp, vp, vp, vp, C.c_uint64, C.c_int,
                                       C.POINTER(Outputs), vp, C.c_char_p, C.c_size_t]
        L.fdt_set_params.argtypes = [vp, _FP, _FP]
        L.fdt_mcmc.argtypes = [vp, C.c_uint64]
        L.fdt_finish.argtypes = [vp]
        L.fdt_get_outputs.argtypes = [vp, C.POINTER(Outputs)]
        L.fdt_run.argtypes = [vp, C.c_uint64]
        L.fdt_set_grid.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_longlong)]
        L.fdt_get_spatial_state.argtypes = [vp, ip, _FP, _FP, _FP, C.POINTER(C.c_double)]
        L.fdt_sync.argtypes = [vp]
        L.fdt_get_timing.argtypes = [vp, C.POINTER(Timing)]
        L.fdt_set_mcmc_tile.argtypes = [vp, C.c_int]
        L.fdt_queue_create.argtypes = [C.c_int]
        L.fdt_queue_create.restype = vp
        L.fdt_queue_destroy.argtypes = [vp]
        L.fdt_queue_destroy.restype = None
        L.fdt_queue_last_error.argtypes = [vp]
        L.fdt_queue_last_error.restype = C.c_char_p
        L.fdt_enqueue_batch.argtypes = [vp, C.POINTER(Opts), C.c_int, _FP, _FP, _FP, C.c_int, C.c_int, C.c_uint64]
        L.fdt_run_queue.argtypes = [vp]
        L.fdt_queue_num_jobs.argtypes = [vp]
        L.fdt_queue_get_outputs.argtypes = [vp, C.c_int, C.POINTER(Outputs)]
        L.fdt_queue_get_params.argtypes = [vp, C.c_int, _FP, _FP]
        L.fdt_queue_set_tiles.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.fdt_queue_set_grad_dev.argtypes = [vp, C.c_int, _FP, C.c_int]
        L.fdt_queue_autotune.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
        L.fdt_queue_timing.argtypes = [vp, C.POINTER(C.c_float), ip, ip, C.c_int]
        L.fdt_get_widths.argtypes = [vp, _FP]
        L.fdt_set_widths.argtypes = [vp, _FP]
        L.fdt_get_energy.argtypes = [vp, _FP]
        L.fdt_surrogate_create.argtypes = [C.c_int, C.c_int, C.c_uint64]
        L.fdt_surrogate_create.restype = vp
        L.fdt_surrogate_destroy.argtypes = [vp]
        L.fdt_surrogate_destroy.restype = None
        L.fdt_surrogate_last_error.argtypes = [vp]
        L.fdt_surrogate_last_error.restype = C.c_char_p
        L.fdt_surrogate_num_weights.argtypes = [vp]
        L.fdt_surrogate_num_weights.restype = C.c_size_t
        L.fdt_surrogate_load.argtypes = [vp, _FP, C.c_size_t]
        L.fdt_surrogate_save.argtypes = [vp, _FP, C.c_size_t]
        L.fdt_surrogate_set_training_data.argtypes = [vp, _FP, _FP, _FP, C.c_int]
        L.fdt_surrogate_gradients.argtypes = [vp, C.c_int, C.c_int, _FP, _FP]
        L.fdt_surrogate_train_step.argtypes = [vp, C.c_int, C.c_int, C.c_float, _FP]
        L.fdt_surrogate_forward.argtypes = [vp, _FP, _FP, C.c_int, _FP]
        L.fdt_surrogate_infer.argtypes = [vp, vp]
        L.fdt_run_surrogate.argtypes = [vp, vp, C.c_uint64, C.c_int, C.c_float, C.c_float, ip]
        L.fdt_get_fallback_flags.argtypes = [vp, C.POINTER(C.c_uint8)]
        _lib = L
    return _lib


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().fdt_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_FP)


class FdtError(RuntimeError):
    pass


_hip_rt = None


def _hip():
    """the HIP runtime the library is linked against (for page-locked host memory)"""
    global _hip_rt
    if _hip_rt is None:
        lib()
        h = None
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                h = C.CDLL(name)
                break
            except OSError:
                continue
        if h is None:
            raise FdtError("the HIP runtime (libamdhip64.so) is not loadable")
        h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        h.hipHostMalloc.restype = C.c_int
        h.hipHostFree.argtypes = [C.c_void_p]
        h.hipHostFree.restype = C.c_int
        _hip_rt = h
    return _hip_rt


def pinned_empty(shape, dtype=np.float32):
    """A numpy array in page-locked host memory (hipHostMalloc): load_data from
    it and outputs(into=...) into it copy at the DMA rate instead of through the
    runtime's staging of pageable memory.  Freed when the array and its views
    are gone (memory still held at interpreter exit is left to the process end)."""
    dt = np.dtype(dtype)
    count = int(np.prod(shape))
    nbytes = max(count * dt.itemsize, 1)
    h = _hip()
    p = C.c_void_p()
    rc = h.hipHostMalloc(C.byref(p), nbytes, 0)
    if rc != 0 or not p.value:
        raise FdtError(f"hipHostMalloc of {nbytes} bytes failed (error {rc})")
    buf = (C.c_char * nbytes).from_address(p.value)
    fin = weakref.finalize(buf, h.hipHostFree, p.value)
    fin.atexit = False
    return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)


class Xfibres:
    """One engine context on one GPU (mirrors the reference's per-part
    xfibres_gpu() call, CUDA/xfibres_gpu.cu:29-146, split into phases)."""

    def __init__(self, opts: Opts | None = None, device: int = 0, **kw):
        self.L = lib()
        self.o = opts if opts is not None else default_opts(**kw)
        self.h = self.L.fdt_create(device, C.byref(self.o))
        if not self.h:
            raise FdtError("fdt_create failed (bad options or no HIP device)")
        self.N = 0
        self.nvox = 0

    def close(self):
        if self.h:
            self.L.fdt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise FdtError(f"fdt error {rc}: {self.L.fdt_last_error(self.h).decode()}")

    @property
    def K(self):
        return self.o.nfibres

    @property
    def nsamples(self):
        return self.o.njumps // self.o.sampleevery

    def set_acquisition(self, bvals, bvecs):
        b, bp = _f32(bvals)
        g, gp = _f32(np.asarray(bvecs).reshape(3, -1))
        self._chk(self.L.fdt_set_acquisition(self.h, b.size, bp, gp))
        self.N = b.size

    def load_data(self, data, layout=LAYOUT_VOX_MAJOR):
        d, dp = _f32(data)
        nvox = d.shape[0] if layout == LAYOUT_VOX_MAJOR else d.shape[1]
        self._chk(self.L.fdt_load_data(self.h, dp, layout, nvox))
        self.nvox = nvox

    def load_grad_dev(self, grad, layout=LAYOUT_VOX_MAJOR):
        """--gradnonlin: per-voxel gradient deviations, [nvox][9] (or [9][nvox]
        with LAYOUT_VOL_MAJOR); None returns to the shared acquisition"""
        if grad is None:
            self._chk(self.L.fdt_load_grad_dev(self.h, _FP(), 0, 0))
            return
        g, gp = _f32(grad)
        nvox = g.shape[0] if layout == LAYOUT_VOX_MAJOR else g.shape[1]
        self._chk(self.L.fdt_load_grad_dev(self.h, gp, layout, nvox))

    def voxel_acquisition(self, vox):
        """(bvals [N], bvecs [3][N]) that voxel `vox` is fitted with"""
        b = np.zeros(self.N, np.float32)
        g = np.zeros((3, self.N), np.float32)
        self._chk(self.L.fdt_get_voxel_acquisition(self.h, int(vox), b.ctypes.data_as(_FP), g.ctypes.data_as(_FP)))
        return b, g

    @property
    def num_params(self):
        return self.L.fdt_num_params(self.h)

    def fit(self):
        self._chk(self.L.fdt_fit(self.h))

    def lm_stage(self, kind, K, f0, G, R, start):
        np_ = (3 if kind == KIND_MULTI else 2) + 3 * K + (1 if f0 else 0)
        s, sp = _f32(np.asarray(start).reshape(self.nvox, np_))
        out = np.zeros((self.nvox, np_), np.float32)
        it = np.zeros(self.nvox, np.int32)
        self._chk(self.L.fdt_lm_stage(self.h, kind, K, int(f0), G, float(R), sp, out.ctypes.data_as(_FP),
                                      it.ctypes.data_as(C.POINTER(C.c_int))))
        return out, it

    def lm_start(self, kind, K, f0):
        np_ = 2 + 3 * K + (1 if f0 else 0)
        out = np.zeros((self.nvox, np_), np.float32)
        self._chk(self.L.fdt_lm_start(self.h, kind, K, int(f0), out.ctypes.data_as(_FP)))
        return out

    def get_params(self):
        p = np.zeros((self.nvox, self.num_params), np.float32)
        t = np.zeros(self.nvox, np.float32)
        self._chk(self.L.fdt_get_params(self.h, p.ctypes.data_as(_FP), t.ctypes.data_as(_FP)))
        return p, t

    def lm_iters(self):
        """LM loop trips per voxel of the last fit (all stages)"""
        it = np.zeros(self.nvox, np.int32)
        self._chk(self.L.fdt_get_lm_iters(self.h, it.ctypes.data_as(C.POINTER(C.c_int))))
        return it

    def set_params(self, params, tau=None):
        p, pp = _f32(params)
        tp = _f32(tau)[1] if tau is not None else _FP()
        self._chk(self.L.fdt_set_params(self.h, pp, tp))

    def mcmc(self, vox_offset=0):
        self._chk(self.L.fdt_mcmc(self.h, vox_offset))

    def finish(self):
        self._chk(self.L.fdt_finish(self.h))

    def run(self, vox_offset=0):
        self._chk(self.L.fdt_run(self.h, vox_offset))

    def sync(self):
        self._chk(self.L.fdt_sync(self.h))

    def set_grid(self, dims, gidx):
        """grid (nx, ny, nz) and each loaded voxel's index x + nx (y + ny z), increasing (mask order)"""
        g = np.ascontiguousarray(gidx, dtype=np.int64)
        self._chk(self.L.fdt_set_grid(self.h, int(dims[0]), int(dims[1]), int(dims[2]),
                                      g.ctypes.data_as(C.POINTER(C.c_longlong))))

    def spatial_state(self):
        """chain-order state a neighbour-initialised run left: been, mean_f [K][nvox], mean_d, mean_S0, dyad [K][3][nvox]"""
        K, nv = self.o.nfibres, self.nvox
        st = dict(been=np.zeros(nv, np.int32), mean_f=np.zeros((K, nv), np.float32), mean_d=np.zeros(nv, np.float32),
                  mean_S0=np.zeros(nv, np.float32), dyad=np.zeros((K, 3, nv), np.float64))
        self._chk(self.L.fdt_get_spatial_state(self.h, st["been"].ctypes.data_as(C.POINTER(C.c_int)),
                                               st["mean_f"].ctypes.data_as(_FP), st["mean_d"].ctypes.data_as(_FP),
                                               st["mean_S0"].ctypes.data_as(_FP),
                                               st["dyad"].ctypes.data_as(C.POINTER(C.c_double))))
        return st

    def timing(self):
        t = Timing()
        self._chk(self.L.fdt_get_timing(self.h, C.byref(t)))
        return dict(ms_fit=t.ms_fit, ms_mcmc=t.ms_mcmc, ms_finish=t.ms_finish, vpl=t.mcmc_vpl, lanes=t.mcmc_lanes)

    def set_mcmc_tile(self, vpl):
        self._chk(self.L.fdt_set_mcmc_tile(self.h, vpl))

    def set_widths(self, widths):
        """initial proposal widths [nvox][16] per slot for the next chains (<= 0: default); None clears"""
        if widths is None:
            self._chk(self.L.fdt_set_widths(self.h, None))
            return
        w, wp = _f32(np.asarray(widths).reshape(self.nvox, 16))
        self._chk(self.L.fdt_set_widths(self.h, wp))

    def energy(self):
        """the likelihood energy each chain carries at its last sweep (Gaussian: the running SSR)"""
        e = np.zeros(self.nvox, np.float32)
        self._chk(self.L.fdt_get_energy(self.h, e.ctypes.data_as(_FP)))
        return e

    def widths(self):
        w = np.zeros((self.nvox, 16), np.float32)
        self._chk(self.L.fdt_get_widths(self.h, w.ctypes.data_as(_FP)))
        return w

    def run_surrogate(self, surrogate, vox_offset=0, nburn_short=200, chi2_max=2.0, ang_max_deg=20.0):
        """learned start + short burn-in, plain fallback; returns the number of rerun voxels"""
        n = C.c_int()
        rc = self.L.fdt_run_surrogate(self.h, surrogate.h, vox_offset, nburn_short, chi2_max, ang_max_deg, C.byref(n))
        if rc < 0:
            raise FdtError(self.L.fdt_last_error(self.h).decode())
        return n.value

    def fallback_flags(self):
        f = np.zeros(self.nvox, np.uint8)
        self._chk(self.L.fdt_get_fallback_flags(self.h, f.ctypes.data_as(C.POINTER(C.c_uint8))))
        return f

    def surrogate_start(self, surrogate):
        surrogate._chk(self.L.fdt_surrogate_infer(surrogate.h, self.h))

    def output_buffers(self, raw_scalars=True, alloc=None):
        """arrays of the shapes outputs() fills, from alloc(shape) (e.g.
        pinned_empty; default zero-filled numpy), to pass as outputs(into=...)"""
        return _alloc_outputs(self.o, self.nvox, raw_scalars, alloc)[0]

    def outputs(self, raw_scalars=True, into=None):
        """every output of the last finished run; into: arrays from
        output_buffers() (same options and voxel count) filled in place"""
        if into is None:
            arr, out = _alloc_outputs(self.o, self.nvox, raw_scalars)
        else:
            want = _alloc_outputs(self.o, self.nvox, raw_scalars, shapes_only=True)
            for k, shape in want.items():
                a = into.get(k)
                if a is None or a.shape != shape or a.dtype != np.float32 or not a.flags.c_contiguous:
                    raise ValueError(f"outputs(into=...): {k} must be a C-contiguous float32 array of shape {shape}")
            arr = into
            out = Outputs()
            for n in _OUT_FIELDS:
                setattr(out, n, arr[n].ctypes.data_as(_FP) if n in want else _FP())
        self._chk(self.L.fdt_get_outputs(self.h, C.byref(out)))
        return arr


def _alloc_outputs(o, nv, raw_scalars=True, alloc=None, shapes_only=False):
    """buffers for every output the options produce (alloc(shape), default
    zero-filled numpy), and the fdt_outputs view of them; shapes_only: the
    shapes alone"""
    K, ns = o.nfibres, o.njumps // o.sampleevery
    shapes = dict(th=(K, ns, nv), ph=(K, ns, nv), f=(K, ns, nv), mean_S0=(nv,), mean_d=(nv,), mean_f=(K, nv),
                  dyads=(K, 3, nv), disp=(K, nv))
    if o.model >= 2:
        shapes["mean_dstd"] = (nv,)
    if o.model == 3:
        shapes["mean_R"] = (nv,)
    if o.f0:
        shapes["mean_f0"] = (nv,)
    if o.rician:
        shapes["mean_tau"] = (nv,)
    if raw_scalars:
        shapes["S0"] = (ns, nv)
        shapes["d"] = (ns, nv)
        if o.model >= 2:
            shapes["dstd"] = (ns, nv)
        if o.model == 3:
            shapes["R"] = (ns, nv)
        if o.f0:
            shapes["f0"] = (ns, nv)
        if o.rician:
            shapes["tau"] = (ns, nv)
    if shapes_only:
        return shapes
    make = alloc or (lambda shape: np.zeros(shape, np.float32))
    arr = {k: make(shape) for k, shape in shapes.items()}
    out = Outputs()
    for n in _OUT_FIELDS:
        setattr(out, n, arr[n].ctypes.data_as(_FP) if n in arr else _FP())
    return arr, out


SG_OUT = 28
# proposal slots whose log widths the network predicts (f0, d, d_std, R, S0, th/ph/f of three fibres)
SG_WIDTH_SLOTS = [0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14]


class Surrogate:
    """The learned start (fdt_surrogate_* of the C ABI): MLP N -> 512 -> 512 ->
    256 -> 28 with BF16 MFMA kernels, trained on the GPU."""

    def __init__(self, n_in: int, seed: int = 1, device: int = 0):
        self.L = lib()
        self.h = self.L.fdt_surrogate_create(device, n_in, seed)
        if not self.h:
            raise FdtError("fdt_surrogate_create failed")
        self.n_in = n_in
        self.nw = self.L.fdt_surrogate_num_weights(self.h)

    def close(self):
        if self.h:
            self.L.fdt_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc < 0:
            raise FdtError(self.L.fdt_surrogate_last_error(self.h).decode())
        return rc

    def weights(self):
        w = np.zeros(self.nw, np.float32)
        self._chk(self.L.fdt_surrogate_save(self.h, w.ctypes.data_as(_FP), self.nw))
        return w

    def load(self, w):
        w, wp = _f32(w)
        self._chk(self.L.fdt_surrogate_load(self.h, wp, w.size))

    def layers(self, w=None):
        """(W, b) per layer from a flat weight vector (W is in x out)"""
        w = self.weights() if w is None else w
        dims = [self.n_in, 512, 512, 256, SG_OUT]
        out, o = [], 0
        for l in range(4):
            W = w[o:o + dims[l] * dims[l + 1]].reshape(dims[l], dims[l + 1])
            o += dims[l] * dims[l + 1]
            b = w[o:o + dims[l + 1]]
            o += dims[l + 1]
            out.append((W, b))
        return out

    def set_training_data(self, signals, bvals, targets):
        x, xp = _f32(signals)
        b, bp = _f32(bvals)
        t, tp = _f32(targets)
        self._chk(self.L.fdt_surrogate_set_training_data(self.h, xp, bp, tp, x.shape[0]))
        self.ntrain = x.shape[0]

    def gradients(self, first, batch):
        g = np.zeros(self.nw, np.float32)
        loss = C.c_float()
        self._chk(self.L.fdt_surrogate_gradients(self.h, first, batch, g.ctypes.data_as(_FP), C.byref(loss)))
        return g, loss.value

    def train_step(self, first, batch, lr):
        loss = C.c_float()
        self._chk(self.L.fdt_surrogate_train_step(self.h, first, batch, lr, C.byref(loss)))
        return loss.value

    def forward(self, signals, bvals):
        x, xp = _f32(signals)
        b, bp = _f32(bvals)
        out = np.zeros((x.shape[0], SG_OUT), np.float32)
        self._chk(self.L.fdt_surrogate_forward(self.h, xp, bp, x.shape[0], out.ctypes.data_as(_FP)))
        return out


class Queue:
    """A batch of volumes with their own acquisitions, models and fibre counts
    run together: concurrent LM stages, one persistent chain launch over all
    volumes, hipGraph replay from the second run on (fdt_*queue* of the C ABI)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        self.h = self.L.fdt_queue_create(device)
        if not self.h:
            raise FdtError("fdt_queue_create failed (no HIP device)")
        self.jobs = []

    def close(self):
        if self.h:
            self.L.fdt_queue_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc < 0:
            raise FdtError(self.L.fdt_queue_last_error(self.h).decode())
        return rc

    def enqueue(self, bvals, bvecs, data, vox_offset=0, layout=LAYOUT_VOX_MAJOR, **kw):
        o = default_opts(**kw)
        bv, bvp = _f32(bvals)
        bc, bcp = _f32(bvecs)
        d, dp = _f32(data)
        N = bv.size
        nvox = d.size // N
        j = self._chk(self.L.fdt_enqueue_batch(self.h, C.byref(o), N, bvp, bcp, dp, layout, nvox, vox_offset))
        self.jobs.append((o, nvox))
        return j

    def run(self):
        self._chk(self.L.fdt_run_queue(self.h))

    def set_grad_dev(self, job, grad, layout=LAYOUT_VOX_MAJOR):
        """--gradnonlin for one job: deviations [nvox][9] (None clears)"""
        if grad is None:
            self._chk(self.L.fdt_queue_set_grad_dev(self.h, job, _FP(), layout))
        else:
            g, gp = _f32(grad)
            self._chk(self.L.fdt_queue_set_grad_dev(self.h, job, gp, layout))

    def set_tiles(self, rows_per_wg=0, job=-1, vpl=0):
        self._chk(self.L.fdt_queue_set_tiles(self.h, rows_per_wg, job, vpl))

    def autotune(self, jumps=50, sample_vox=4096):
        ms = C.c_float()
        self._chk(self.L.fdt_queue_autotune(self.h, jumps, sample_vox, C.byref(ms)))
        return ms.value

    def timing(self):
        ms = C.c_float()
        rows = C.c_int()
        n = len(self.jobs)
        vpl = (C.c_int * max(n, 1))()
        self._chk(self.L.fdt_queue_timing(self.h, C.byref(ms), C.byref(rows), vpl, n))
        return dict(ms=ms.value, rows_per_wg=rows.value, vpl=[vpl[i] for i in range(n)])

    def params(self, job):
        o, nvox = self.jobs[job]
        np_ = 2 + 3 * o.nfibres + (1 if o.model >= 2 else 0) + (1 if o.f0 else 0)
        p = np.zeros((nvox, np_), np.float32)
        t = np.zeros(nvox, np.float32)
        self._chk(self.L.fdt_queue_get_params(self.h, job, p.ctypes.data_as(_FP), t.ctypes.data_as(_FP)))
        return p, t

    def outputs(self, job, raw_scalars=True):
        o, nvox = self.jobs[job]
        arr, out = _alloc_outputs(o, nvox, raw_scalars)
        self._chk(self.L.fdt_queue_get_outputs(self.h, job, C.byref(out)))
        return arr


def predict(bvals, bvecs, S0, d, f, dyads, model=1, dstd=None, R=None, f0=None, device=0):
    """posterior-predictive signal [N][nvox] from the posterior means (fdt_predict)"""
    bv, bvp = _f32(bvals)
    bc, bcp = _f32(np.asarray(bvecs).reshape(3, -1))
    S0, S0p = _f32(S0)
    d, dp = _f32(d)
    f, fp = _f32(np.asarray(f).reshape(-1, S0.size) if S0.size else np.zeros((np.asarray(f).shape[0], 0)))
    dy, dyp = _f32(np.asarray(dyads).reshape(f.shape[0], 3, S0.size))
    opt = lambda a: _f32(a)[1] if a is not None else _FP()
    out = np.zeros((bv.size, S0.size), np.float32)
    rc = lib().fdt_predict(device, bv.size, bvp, bcp, S0.size, f.shape[0], model, S0p, dp, opt(dstd), opt(R), opt(f0),
                           fp, dyp, out.ctypes.data_as(_FP))
    if rc:
        raise FdtError(f"fdt_predict failed ({rc})")
    return out


def run_xfibres(bvals, bvecs, data, seed=8665904, vox_offset=0, layout=LAYOUT_VOX_MAJOR, grad=None, **opts):
    """Whole pipeline on one GPU: LM init -> MCMC -> finish; returns outputs and
    params.  grad: per-voxel gradient deviations [nvox][9] (--gradnonlin)."""
    x = Xfibres(seed=seed, **opts)
    try:
        x.set_acquisition(bvals, bvecs)
        x.load_data(data, layout)
        if grad is not None:
            x.load_grad_dev(grad)
        x.fit()
        params, tau = x.get_params()
        x.mcmc(vox_offset)
        x.finish()
        out = x.outputs()
        out["params"] = params
        out["tau_init"] = tau
        return out
    finally:
        x.close()
