"""Writes the golden fixtures of the FP64 oracle (oracle/xfibres_oracle.cpp).

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_golden.py
The fixtures are the oracle's outputs on small seeded synthetic inputs; the
tests check that the oracle still reproduces them (a guard against drift of
the restatement) and that the HIP engine lands on them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fdt_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    out = {}
    # LM stages: (kind, G, K, f0) on 8 voxels each
    for name, kind, G, K, f0 in [("c0", 0, 0, 2, 0), ("c0", 1, 0, 2, 1), ("c1", 2, 0, 3, 0), ("c1", 2, 2, 2, 0)]:
        v = synth.voxels(name, 8)
        y = v["data"].astype(np.float64)
        if kind < 2:
            st = O.start(kind, K, bool(f0), v["bvals"], v["bvecs"], y)
        else:
            s0 = O.start(O.K_SINGLE, K, bool(f0), v["bvals"], v["bvecs"], y)
            s1, _, _ = O.lm(O.K_SINGLE, K, bool(f0), 0, 0.0, v["bvals"], v["bvecs"], y, s0)
            st = np.zeros((8, 3 + 3 * K + f0))
            st[:, 0], st[:, 1], st[:, 2] = s1[:, 0], 1.0, s1[:, 1]
            for k in range(K):
                st[:, 3 + 3 * k] = np.tan(s1[:, 2 + 3 * k] / 0.636619772)
                st[:, 4 + 3 * k] = s1[:, 3 + 3 * k]
                st[:, 5 + 3 * k] = s1[:, 4 + 3 * k]
        st = st.astype(np.float32).astype(np.float64)
        res, it, cost = O.lm(kind, K, bool(f0), G, 0.13, v["bvals"], v["bvecs"], y, st)
        key = f"lm_{name}_k{kind}_g{G}_n{K}_f{f0}"
        out[key + "_data"] = v["data"]
        out[key + "_start"] = st
        out[key + "_out"] = res
        out[key + "_iters"] = it
        out[key + "_cost"] = cost
    # full fit pipeline
    for name, model, K, extra in [("c0", 1, 2, {}), ("c1", 2, 3, {}), ("c1", 3, 2, {"rician": 1})]:
        v = synth.voxels(name, 8)
        o = O.opts(nfibres=K, model=model, **extra)
        p, tau, it = O.fit(o, v["bvals"], v["bvecs"], v["data"].astype(np.float64))
        key = f"fit_{name}_m{model}_n{K}" + ("_ric" if extra else "")
        out[key + "_data"] = v["data"]
        out[key + "_params"] = p
        out[key + "_tau"] = tau
    # MCMC: 8 voxels, 100 jumps, from the oracle's fit
    v = synth.voxels("c0", 8)
    o = O.opts(nfibres=2, model=1, nburn=50, njumps=100, sampleevery=25)
    p, tau, _ = O.fit(O.opts(nfibres=2, model=1), v["bvals"], v["bvecs"], v["data"].astype(np.float64))
    p = p.astype(np.float32).astype(np.float64)
    tau = tau.astype(np.float32).astype(np.float64)
    s = O.mcmc(o, v["bvals"], v["bvecs"], v["data"].astype(np.float64), p, tau)
    out["mcmc_c0_data"] = v["data"]
    out["mcmc_c0_params"] = p
    out["mcmc_c0_tau"] = tau
    for k in ("th", "ph", "f", "S0", "d", "energy"):
        out[f"mcmc_c0_{k}"] = np.asarray(s[k], dtype=np.float64)
    for name in ("c0", "c1"):
        v = synth.voxels(name, 1)
        out[f"acq_{name}_bvals"] = v["bvals"]
        out[f"acq_{name}_bvecs"] = v["bvecs"]
    np.savez_compressed(os.path.join(HERE, "oracle_golden.npz"), **out)
    print(f"wrote {len(out)} arrays to tests/golden/oracle_golden.npz")


if __name__ == "__main__":
    main()
