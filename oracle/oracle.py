"""ctypes binding of the oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  orc64_* is the FP64
oracle, orc32_* the FP32 restatement timed as the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

K_SINGLE, K_SINGLE_C, K_MULTI = 0, 1, 2


class OrcOpts(C.Structure):
    _fields_ = [
        ("nfibres", C.c_int), ("model", C.c_int), ("fudge", C.c_double), ("njumps", C.c_int),
        ("nburn", C.c_int), ("sampleevery", C.c_int), ("updateproposalevery", C.c_int),
        ("seed", C.c_uint64), ("no_ard", C.c_int), ("all_ard", C.c_int), ("nonlin", C.c_int),
        ("cnonlin", C.c_int), ("rician", C.c_int), ("f0", C.c_int), ("ardf0", C.c_int),
        ("Rmean", C.c_double), ("Rstd", C.c_double), ("Rfudge", C.c_double),
    ]


class OrcSamples(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n in ("th", "ph", "f", "S0", "d", "dstd", "R", "f0", "tau")]


def opts(**kw) -> OrcOpts:
    d = dict(nfibres=1, model=1, fudge=1.0, njumps=1250, nburn=1000, sampleevery=25, updateproposalevery=40,
             seed=8665904, no_ard=0, all_ard=0, nonlin=0, cnonlin=1, rician=0, f0=0, ardf0=0,
             Rmean=0.13, Rstd=0.03, Rfudge=0.0)
    d.update(kw)
    return OrcOpts(**d)


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        for p in ("orc64", "orc32"):
            getattr(L, p + "_lm").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, dp, dp,
                                             C.c_int, dp, dp, dp, ip, dp]
            getattr(L, p + "_model_eval").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, dp,
                                                     dp, dp, dp, dp, dp, dp, dp]
            getattr(L, p + "_lu_solve").argtypes = [C.c_int, dp, dp, dp]
            getattr(L, p + "_start").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, C.c_int, dp, dp, dp]
            getattr(L, p + "_fit").argtypes = [C.POINTER(OrcOpts), C.c_int, dp, dp, C.c_int, dp, dp, dp, ip, dp]
            getattr(L, p + "_mcmc").argtypes = [C.POINTER(OrcOpts), C.c_int, dp, dp, C.c_int, dp, dp, dp,
                                               C.c_uint64, C.POINTER(OrcSamples), dp, dp, dp]
        L.orc_dti.argtypes = [C.c_int, dp, dp, C.c_int, dp, dp]
        L.orc_correct_acq.argtypes = [C.c_int, dp, dp, dp, dp, dp]
        L.orc_pinv.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_eig.argtypes = [C.c_int, dp, dp, dp]
        u32p = C.POINTER(C.c_uint32)
        L.orc_philox.argtypes = [u32p, u32p, u32p]
        L.orc_draw.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, dp, dp]
        L.orc_logI0.argtypes = [C.c_double]
        L.orc_logI0.restype = C.c_double
        L.orc_mcmc_predict.argtypes = [C.POINTER(OrcOpts), C.c_int, dp, dp, dp, dp]
        L.orc_finish.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OrcSamples)] + [dp] * 9
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def n_params(kind: int, K: int, f0: bool) -> int:
    return (3 if kind == K_MULTI else 2) + 3 * K + (1 if f0 else 0)


def model_eval(kind, K, f0, G, Rv, bvals, bvecs, p, y=None, prec=64):
    L = lib()
    N = len(bvals)
    np_ = n_params(kind, K, f0)
    bv, bvp = _d(bvals)
    bc, bcp = _d(bvecs)
    pp, ppp = _d(p)
    pred = np.zeros(N)
    Jg = np.zeros((N, np_))
    Jh = np.zeros((N, np_))
    cost = C.c_double(0)
    yy, yyp = _d(y if y is not None else np.zeros(N))
    getattr(L, f"orc{prec}_model_eval")(kind, K, int(f0), G, Rv, N, bvp, bcp, ppp,
                                        pred.ctypes.data_as(C.POINTER(C.c_double)),
                                        Jg.ctypes.data_as(C.POINTER(C.c_double)),
                                        Jh.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cost), yyp)
    return pred, Jg, Jh, cost.value


def lm(kind, K, f0, G, Rv, bvals, bvecs, data, start, prec=64):
    L = lib()
    data = np.ascontiguousarray(data, dtype=np.float64)
    nvox, N = data.shape
    np_ = n_params(kind, K, f0)
    start = np.ascontiguousarray(start, dtype=np.float64).reshape(nvox, np_)
    out = np.zeros((nvox, np_))
    iters = np.zeros(nvox, dtype=np.int32)
    cost = np.zeros(nvox)
    getattr(L, f"orc{prec}_lm")(kind, K, int(f0), G, Rv, N, _d(bvals)[1], _d(bvecs)[1], nvox, _d(data)[1],
                                _d(start)[1], out.ctypes.data_as(C.POINTER(C.c_double)),
                                iters.ctypes.data_as(C.POINTER(C.c_int)), cost.ctypes.data_as(C.POINTER(C.c_double)))
    return out, iters, cost


def _grad(grad, nvox):
    """per-voxel gradient-deviation values [nvox][9] (--gradnonlin) or a null pointer"""
    if grad is None:
        return None, C.POINTER(C.c_double)()
    g = np.ascontiguousarray(grad, dtype=np.float64).reshape(nvox, 9)
    return g, g.ctypes.data_as(C.POINTER(C.c_double))


def correct_acq(bvals, bvecs, g):
    """corrected (bvals [N], bvecs [3][N]) of one voxel with gradient deviation g [9]"""
    N = len(bvals)
    bv_c = np.zeros(N)
    bc_c = np.zeros((3, N))
    lib().orc_correct_acq(N, _d(bvals)[1], _d(bvecs)[1], _d(g)[1], bv_c.ctypes.data_as(C.POINTER(C.c_double)),
                          bc_c.ctypes.data_as(C.POINTER(C.c_double)))
    return bv_c, bc_c


def start(kind, K, f0, bvals, bvecs, data, prec=64, grad=None):
    L = lib()
    data = np.ascontiguousarray(data, dtype=np.float64)
    nvox, N = data.shape
    np_ = n_params(kind, K, f0)
    out = np.zeros((nvox, np_))
    bv, bc = _d(bvals), _d(bvecs)
    g = _grad(grad, nvox)
    getattr(L, f"orc{prec}_start")(kind, K, int(f0), N, bv[1], bc[1], nvox, _d(data)[1],
                                   out.ctypes.data_as(C.POINTER(C.c_double)), g[1])
    return out


def lu_solve(A, b, prec=64):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    x = np.zeros(n)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    getattr(lib(), f"orc{prec}_lu_solve")(n, A.ctypes.data_as(C.POINTER(C.c_double)),
                                          bb.ctypes.data_as(C.POINTER(C.c_double)),
                                          x.ctypes.data_as(C.POINTER(C.c_double)))
    return x


def dti(bvals, bvecs, data):
    data = np.ascontiguousarray(data, dtype=np.float64).reshape(-1, len(bvals))
    out = np.zeros((data.shape[0], 15))
    bv, bc = _d(bvals), _d(bvecs)
    lib().orc_dti(len(bvals), bv[1], bc[1], data.shape[0], data.ctypes.data_as(C.POINTER(C.c_double)),
                  out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def pinv(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    m, n = A.shape
    P = np.zeros((n, m))
    lib().orc_pinv(m, n, A.ctypes.data_as(C.POINTER(C.c_double)), P.ctypes.data_as(C.POINTER(C.c_double)))
    return P


def eig(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_eig(n, A.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)),
                  V.ctypes.data_as(C.POINTER(C.c_double)))
    return w, V


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return [int(v) for v in o]


def draw(seed, vox, jump, slot):
    n, lu = C.c_double(), C.c_double()
    lib().orc_draw(seed, vox, jump, slot, C.byref(n), C.byref(lu))
    return n.value, lu.value


def logI0(x):
    return lib().orc_logI0(float(x))


def fit(o: OrcOpts, bvals, bvecs, data, prec=64, grad=None):
    data = np.ascontiguousarray(data, dtype=np.float64)
    nvox, N = data.shape
    npo = 2 + 3 * o.nfibres + (1 if o.model >= 2 else 0) + (1 if o.f0 else 0)
    params = np.zeros((nvox, npo))
    tau = np.zeros(nvox)
    iters = np.zeros(nvox, dtype=np.int32)
    bv, bc = _d(bvals), _d(bvecs)
    getattr(lib(), f"orc{prec}_fit")(C.byref(o), N, bv[1], bc[1], nvox, data.ctypes.data_as(C.POINTER(C.c_double)),
                                     params.ctypes.data_as(C.POINTER(C.c_double)),
                                     tau.ctypes.data_as(C.POINTER(C.c_double)),
                                     iters.ctypes.data_as(C.POINTER(C.c_int)), _grad(grad, nvox)[1])
    return params, tau, iters


def alloc_samples(o: OrcOpts, nvox: int):
    K, ns = o.nfibres, o.njumps // o.sampleevery
    arrs = dict(th=np.zeros((K, ns, nvox), np.float32), ph=np.zeros((K, ns, nvox), np.float32),
                f=np.zeros((K, ns, nvox), np.float32), S0=np.zeros((ns, nvox), np.float32),
                d=np.zeros((ns, nvox), np.float32))
    arrs["dstd"] = np.zeros((ns, nvox), np.float32) if o.model >= 2 else None
    arrs["R"] = np.zeros((ns, nvox), np.float32) if o.model == 3 else None
    arrs["f0"] = np.zeros((ns, nvox), np.float32) if o.f0 else None
    arrs["tau"] = np.zeros((ns, nvox), np.float32) if o.rician else None
    s = OrcSamples()
    for k, v in arrs.items():
        setattr(s, k, v.ctypes.data_as(C.POINTER(C.c_float)) if v is not None else C.POINTER(C.c_float)())
    return arrs, s


def mcmc(o: OrcOpts, bvals, bvecs, data, params, tau=None, vox_offset=0, prec=64, grad=None, widths=None):
    """chains from the start vectors `params`; widths [nvox][16]: initial
    proposal widths per draw slot (<= 0 keeps the default), or None"""
    data = np.ascontiguousarray(data, dtype=np.float64)
    nvox, N = data.shape
    params = np.ascontiguousarray(params, dtype=np.float64)
    arrs, s = alloc_samples(o, nvox)
    energy = np.zeros(nvox)
    tp = _d(tau)[1] if tau is not None else C.POINTER(C.c_double)()
    bv, bc = _d(bvals), _d(bvecs)
    getattr(lib(), f"orc{prec}_mcmc")(C.byref(o), N, bv[1], bc[1], nvox, data.ctypes.data_as(C.POINTER(C.c_double)),
                                      params.ctypes.data_as(C.POINTER(C.c_double)), tp, vox_offset, C.byref(s),
                                      energy.ctypes.data_as(C.POINTER(C.c_double)), _grad(grad, nvox)[1],
                                      _d(widths)[1] if widths is not None else C.POINTER(C.c_double)())
    arrs["energy"] = energy
    return arrs


def finish(o: OrcOpts, arrs):
    K, ns = o.nfibres, o.njumps // o.sampleevery
    nvox = arrs["S0"].shape[1]
    s = OrcSamples()
    for k in ("th", "ph", "f", "S0", "d", "dstd", "R", "f0", "tau"):
        v = arrs.get(k)
        setattr(s, k, v.ctypes.data_as(C.POINTER(C.c_float)) if v is not None else C.POINTER(C.c_float)())
    outs = {k: np.zeros(nvox) for k in ("mean_S0", "mean_d", "mean_dstd", "mean_R", "mean_f0", "mean_tau")}
    outs["mean_f"] = np.zeros((K, nvox))
    outs["dyads"] = np.zeros((K, 3, nvox))
    outs["disp"] = np.zeros((K, nvox))
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    lib().orc_finish(K, ns, nvox, C.byref(s), ptr(outs["mean_S0"]), ptr(outs["mean_d"]), ptr(outs["mean_dstd"]),
                     ptr(outs["mean_R"]), ptr(outs["mean_f0"]), ptr(outs["mean_tau"]), ptr(outs["mean_f"]),
                     ptr(outs["dyads"]), ptr(outs["disp"]))
    return outs


def mcmc_predict(o: OrcOpts, bvals, bvecs, state):
    N = len(bvals)
    pred = np.zeros(N)
    st = np.ascontiguousarray(state, dtype=np.float64)
    bv, bc = _d(bvals), _d(bvecs)
    lib().orc_mcmc_predict(C.byref(o), N, bv[1], bc[1], st.ctypes.data_as(C.POINTER(C.c_double)),
                           pred.ctypes.data_as(C.POINTER(C.c_double)))
    return pred


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))
