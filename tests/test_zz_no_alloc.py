"""fdt_mcmc_noalloc -- the chains and the finishing pass on caller-owned
device buffers, with the library's scratch in a static device array (no
allocation) -- against the context path fdt_set_params + fdt_mcmc +
fdt_finish + fdt_get_outputs on the same inputs: identical outputs, on two
calls in a row, the second with fewer voxels (the scratch keeps the first
call's contents).  The test's own device buffers and stream come from the
HIP runtime the library is linked against; every engine call goes through
the C ABI."""
import ctypes as C

import numpy as np
import pytest

from fdt_amd import synth

pytestmark = pytest.mark.gpu


def context_run(fdt, v, opts):
    x = fdt.Xfibres(**opts)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        x.fit()
        params, tau = x.get_params()
        x.mcmc(0)
        x.finish()
        return params, tau, x.outputs()
    finally:
        x.close()


class Hip:
    """the HIP runtime the library is linked against (already loaded with it)"""

    def __init__(self, fdt):
        fdt.lib()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.live = []

    def chk(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        self.chk(self.h.hipMalloc(C.byref(p), max(a.nbytes, 4)))
        self.live.append(p)
        self.chk(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))   # host to device
        return p

    def download(self, p, shape):
        a = np.empty(shape, np.float32)
        self.chk(self.h.hipMemcpy(a.ctypes.data, p, a.nbytes, 2))   # device to host
        return a

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def noalloc_run(fdt, hip, v, opts, params, tau, want_keys):
    o = fdt.default_opts(**opts)
    N = v["bvals"].size
    nvox = v["data"].shape[0]
    bv, bc, dd, pp = hip.upload(v["bvals"]), hip.upload(v["bvecs"]), hip.upload(v["data"]), hip.upload(params)
    tt = hip.upload(tau) if opts.get("rician") else None
    K, ns = o.nfibres, o.njumps // o.sampleevery
    shapes = {k: (K, ns, nvox) for k in ("th", "ph", "f")}
    shapes.update(mean_f=(K, nvox), dyads=(K, 3, nvox), disp=(K, nvox))
    shapes.update({k: (ns, nvox) for k in ("S0", "d", "dstd", "R", "f0", "tau")})
    outs = {k: hip.upload(np.full(shapes.get(k, (nvox,)), np.nan, np.float32)) for k in want_keys}
    desc = fdt.Outputs()
    for k, p in outs.items():
        setattr(desc, k, C.cast(p, C.POINTER(C.c_float)))
    msg = C.create_string_buffer(256)
    stream = C.c_void_p()
    hip.chk(hip.h.hipStreamCreate(C.byref(stream)))
    try:
        rc = fdt.lib().fdt_mcmc_noalloc(C.byref(o), N, bv, bc, dd, pp, tt, 0, nvox, C.byref(desc), stream, msg, 256)
        assert rc == 0, msg.value.decode()
        hip.chk(hip.h.hipStreamSynchronize(stream))
    finally:
        hip.h.hipStreamDestroy(stream)
    return {k: hip.download(p, shapes.get(k, (nvox,))) for k, p in outs.items()}


@pytest.mark.parametrize("name,model,K,rician", [("c1", 2, 3, 0), ("c1", 3, 2, 1)])
def test_noalloc_matches_context_path(fdt, name, model, K, rician):
    hip = Hip(fdt)
    opts = dict(nfibres=K, model=model, rician=rician, nburn=100, njumps=250, sampleevery=25)
    try:
        for nvox in (300, 77):  # two calls in a row, the second smaller
            v = synth.voxels(name, nvox, sub=nvox)
            params, tau, want = context_run(fdt, v, opts)
            got = noalloc_run(fdt, hip, v, opts, params, tau, want.keys())
            for k, w in want.items():
                np.testing.assert_array_equal(got[k], w, err_msg=f"{nvox} voxels: {k}")
    finally:
        hip.free()


def test_noalloc_rejects_oversize_and_null(fdt):
    o = fdt.default_opts(nfibres=3, model=2)          # 50 samples: 450 floats of scratch per voxel
    msg = C.create_string_buffer(256)
    desc = fdt.Outputs()
    too_many = (4 << 20) // 450 + 1
    rc = fdt.lib().fdt_mcmc_noalloc(C.byref(o), 288, None, None, None, None, None, 0, too_many, C.byref(desc), None,
                                    msg, 256)
    assert rc == -1 and b"scratch" in msg.value
    rc = fdt.lib().fdt_mcmc_noalloc(C.byref(o), 288, None, None, None, None, None, 0, 10, C.byref(desc), None, msg, 256)
    assert rc == -1 and b"null" in msg.value
