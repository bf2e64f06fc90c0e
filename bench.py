"""Throughput of the MI355X xfibres engine on the multi-shell 3-fibre workload.

One step = one pass of the hot path over the whole synthetic volume with the
inputs resident in HBM: LM initialisation (tensor start values + PVM_single +
PVM_multi), 1000 burn-in + 1250 recorded MCMC sweeps, and the finishing pass
(fibre sort, means, dyads).  Workload: model 2 (sticks with gamma-distributed
diffusivities), nf=3, 18 b=0 + 90 directions at each of b=1000/2000/3000
(N=288), an ellipsoid mask of 250,968 voxels in a 110x110x72 grid, Rician
SNR 25, synthetic (seeded) data.

Also reported: the roofline position of the MCMC kernel (algorithmic FP32
flops per launch / its HIP-event duration, against 157.3 TFLOP/s), the
end-to-end rate from host buffers (upload + run + download of all outputs),
and the CPU baseline: the oracle's FP32 restatement of the reference CPU
xfibres, timed on this host on a bounded sample of the same voxels, on all
threads and on one.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FP32_PEAK_TFLOPS = 157.3
# HBM bytes per launch of mcmc_kernel<3,18,false> on the full 250,968-voxel
# workload: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by the same
# counter's reading for a copy kernel of known size (profiles/mcmc_c1_traffic.md)
MCMC_TRAFFIC_BYTES = (2 * 148124 + 632652) * 1024
METRIC = "voxels/s, 3-fibre multi-shell bedpostx; surrogate vs plain at equal dyad error"
LM_MODES = {0: "FP32 model, Jacobian and cost", 1: "FP64 throughout",
            2: "FP32 Jacobian and normal equations, FP64 parameters and cost"}
LM_DEFAULT = 2


def lm_flops_per_iter(N: int, npar: int, K: int = 3) -> int:
    """Algorithmic FP32 flops of one LM loop trip of one voxel: the normal
    equations N*(np(np+1) + 2np) plus ~40K per volume for the model rows."""
    return N * (npar * (npar + 1) + 2 * npar + 40 * K)


def flops_per_jump(model: int, K: int, N: int, f0: bool = False, rician: bool = False) -> int:
    """Algorithmic FP32 flops of one MCMC sweep of one voxel.

    F = N * (P*L + R_sig): P proposals each re-evaluating the likelihood
    (L = 2K+7 per volume, Gaussian; 2K+33 Rician) plus the signal updates
    (d: K*A + I; d_std: K*A + I for model 2, I for model 3; R: K*A; 2K angle
    proposals: 6 + A each), with A = 3/7/5 and I = 2/6/6 for models 1/2/3.
    FMA = 2, add/mul = 1, each exp/log/div/sqrt = 1."""
    P = 3 * K + 2 + (model >= 2) + (model == 3) + int(f0) + int(rician)
    L = 2 * K + (33 if rician else 7)
    A = {1: 3, 2: 7, 3: 5}[model]
    I = {1: 2, 2: 6, 3: 6}[model]
    R = K * A + I
    if model >= 2:
        R += (K * A + I) if model == 2 else I
    if model == 3:
        R += K * A
    R += 2 * K * (6 + A)
    return N * (P * L + R)


def _cpu_rate(O, o, v, threads, seconds, rng, start_n):
    """voxels/s of the FP32 CPU restatement (LM + full chain per voxel) on `threads` threads"""
    O.set_threads(threads)
    n = start_n
    done, t_used = 0, 0.0
    while t_used < seconds and done < 4096:
        idx = rng.choice(v["data"].shape[0], n, replace=False)
        d = v["data"][idx].astype(np.float64)
        t = time.perf_counter()
        p, tau, _ = O.fit(o, v["bvals"], v["bvecs"], d, prec=32)
        O.mcmc(o, v["bvals"], v["bvecs"], d, p, tau, prec=32)
        t_used += time.perf_counter() - t
        done += n
        n = min(2 * n, 4096 - done) if t_used < seconds / 3 else n
    return done / t_used, done, t_used


def host_cpu():
    """CPU model and logical core count of this host"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(v, opts, seconds=15.0, threads=None):
    """The oracle's FP32 restatement of the reference CPU xfibres (LM + full
    chain per voxel), one voxel per OpenMP thread as bedpostx runs parts in
    parallel.  Threads: the host's CPU share for one GPU (16 on the GPU hosts,
    whatever nproc shows), recorded with nproc and the CPU model."""
    from oracle import oracle as O
    threads = threads or min(16, os.cpu_count() or 1)
    o = O.opts(nfibres=opts["nfibres"], model=opts["model"], nburn=opts["nburn"], njumps=opts["njumps"],
               sampleevery=opts["sampleevery"])
    rng = np.random.default_rng(12345)
    rate, done, t_used = _cpu_rate(O, o, v, threads, seconds, rng, max(threads, 16))
    rate1, done1, t1 = _cpu_rate(O, o, v, 1, seconds / 3, rng, 2)
    model, nproc = host_cpu()
    return dict(value=rate, unit="voxels/s", cores=threads, kind="port", value_1core=rate1,
                host_nproc=nproc, host_cpu=model,
                sample=f"{done} voxels drawn at random from the workload, full LM + {opts['nburn']}+{opts['njumps']} "
                       f"MCMC sweeps each, oracle FP32 restatement of xfibres on {threads} OpenMP threads "
                       f"({t_used:.1f} s of CPU work; {threads} is this host's CPU share for one GPU, of {nproc} "
                       f"logical cores); value_1core: {done1} voxels on 1 thread ({t1:.1f} s)")


SURROGATE_WEIGHTS = os.path.join(HERE, "fdt_amd", "weights", "surrogate_c1_model2_nf3.npz")
# learned start on the multi-shell volume: burn-in 200 after the network's start
# as the learned-start workload specifies it, and 50 (the shortest burn-in whose
# median dyad error vs truth stays at the plain path's); the model-3 Rician
# volume keeps burn-in 200
SG_DEFAULTS = dict(nburn_short=200, chi2_max=3.0, ang_max_deg=30.0)
SG_SHORT = dict(nburn_short=50, chi2_max=3.0, ang_max_deg=30.0)
SG_DEFAULTS_C4 = dict(nburn_short=200, chi2_max=3.0, ang_max_deg=30.0)


def dyad_summary(out, v):
    from fdt_amd import surrogate as SG
    e = SG.dyad_errors(out["dyads"], out["mean_f"], v["dirs"], v["f"])
    return {"dyad_err_median_deg": float(np.nanmedian(e)), "unmatched_fibres": float(np.mean(np.isnan(e)))}


def surrogate_eval(v, opts, s, steps, plain_out=None, sg=None):
    """time the learned-start path on the resident volume and compare dyad errors with the plain path"""
    import fdt_amd as F
    sg = sg or SG_DEFAULTS
    nvox = v["data"].shape[0]
    x = F.Xfibres(**opts)
    x.set_acquisition(v["bvals"], v["bvecs"])
    x.load_data(v["data"])
    x.run_surrogate(s, **sg)
    x.sync()
    t = time.perf_counter()
    for _ in range(steps):
        nfb = x.run_surrogate(s, **sg)
    x.sync()
    ms = (time.perf_counter() - t) / steps * 1e3
    out = x.outputs(raw_scalars=False)
    x.close()
    res = {"value": nvox / (ms / 1e3), "unit": "voxels/s", "ms_per_step": ms, "fallback_frac": nfb / nvox,
           **dyad_summary(out, v), **sg}
    if plain_out is not None:
        pl = dyad_summary(plain_out, v)
        res["plain_dyad_err_median_deg"] = pl["dyad_err_median_deg"]
        res["plain_unmatched_fibres"] = pl["unmatched_fibres"]
    return res


def plain_short_burnin(v, opts, nburn):
    """ablation: the plain LM start with the learned path's short burn-in (no network, no fallback)"""
    import fdt_amd as F
    nvox = v["data"].shape[0]
    x = F.Xfibres(**dict(opts, nburn=nburn))
    x.set_acquisition(v["bvals"], v["bvecs"])
    x.load_data(v["data"])
    x.run(0)
    x.sync()
    t = time.perf_counter()
    x.run(0)
    x.sync()
    ms = (time.perf_counter() - t) * 1e3
    out = x.outputs(raw_scalars=False)
    x.close()
    return {"value": nvox / (ms / 1e3), "unit": "voxels/s", "ms_per_step": ms, "nburn": nburn, **dyad_summary(out, v)}


def bench_c3(args):
    """Learned start on the multi-shell volume: train the network from scratch
    on the engine's own labels (plain LM + MCMC on seeded voxels of the same
    acquisition, disjoint from the evaluation voxels), then time the learned
    path (burn-in 200, plain fallback) against the plain path and compare
    median dyad errors against the synthetic truth."""
    import fdt_amd as F
    from fdt_amd import synth, surrogate as SG
    opts = dict(nfibres=3, model=2, nburn=1000, njumps=1250, sampleevery=25)
    t0 = time.perf_counter()
    tr = synth.voxels("c1", args.train_voxels, sub=1, with_truth=False)
    t_syn = time.perf_counter() - t0
    t0 = time.perf_counter()
    T = SG.label(tr["bvals"], tr["bvecs"], tr["data"], **opts)
    t_lab = time.perf_counter() - t0
    s = F.Surrogate(tr["bvals"].size, seed=11)
    t0 = time.perf_counter()
    losses = SG.train(s, tr["data"], tr["bvals"], T, epochs=args.epochs, batch=1024, lr=1e-3, lr_decay=0.8)
    t_tr = time.perf_counter() - t0
    del tr, T
    if args.save_weights:
        os.makedirs(os.path.dirname(os.path.abspath(args.save_weights)), exist_ok=True)
        np.savez_compressed(args.save_weights, weights=s.weights(), n_in=s.n_in, train_voxels=args.train_voxels,
                            epochs=args.epochs, loss=np.array(losses))
    nvox = args.nvox or synth.C1_NVOX
    v = synth.voxels("c1", nvox)
    x = F.Xfibres(**opts)
    x.set_acquisition(v["bvals"], v["bvecs"])
    x.load_data(v["data"])
    x.run(0)
    x.sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        x.run(0)
        x.sync()
    plain_ms = (time.perf_counter() - t) / args.steps * 1e3
    plain_out = x.outputs(raw_scalars=False)
    x.close()
    sg = surrogate_eval(v, opts, s, args.steps, plain_out)
    s.close()
    line = {
        "metric": METRIC, "value": sg["value"], "unit": "voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": sg["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 chains, bf16 network",
        "data": "synthetic (seeded Philox), no published dataset",
        "config": {"workload": f"learned start on multi-shell model 2, nf=3, N=288, {nvox} voxels, "
                               f"burn-in {SG_DEFAULTS['nburn_short']} "
                               "(fallback: plain LM + burn-in 1000), nj=1250 se=25",
                   "plain_voxels_per_s": nvox / (plain_ms / 1e3), "speedup_vs_plain": plain_ms / sg["ms_per_step"],
                   "surrogate": sg,
                   "training": {"voxels": args.train_voxels, "epochs": args.epochs, "final_loss": losses[-1],
                                "s_synth": t_syn, "s_label": t_lab, "s_train": t_tr,
                                "network": "MLP 288-512-512-256-28, GELU, BF16 MFMA, FP32 accumulation, Adam"}},
    }
    print(json.dumps(line), flush=True)


def bench_c4(args):
    """Learned start at high resolution: model 3 (zeppelins) with the Rician
    noise model on a 904,550-voxel brain-shaped volume whose SNR varies
    smoothly in [10, 30] (15% of crossings below 45 deg).  The network is
    retrained on model-3 labels; the plain path runs on the same voxels for comparison; posterior
    f is compared with the FP64 oracle on a random sample of voxels."""
    import fdt_amd as F
    from fdt_amd import synth, surrogate as SG
    from oracle import oracle as O
    opts = dict(nfibres=3, model=3, rician=1, nburn=1000, njumps=1250, sampleevery=25)
    t0 = time.perf_counter()
    tr = synth.voxels("c4", args.train_voxels, sub=1, with_truth=False)
    t_syn = time.perf_counter() - t0
    t0 = time.perf_counter()
    T = SG.label(tr["bvals"], tr["bvecs"], tr["data"], **opts)
    t_lab = time.perf_counter() - t0
    s = F.Surrogate(tr["bvals"].size, seed=13)
    t0 = time.perf_counter()
    losses = SG.train(s, tr["data"], tr["bvals"], T, epochs=args.epochs, batch=1024, lr=1e-3, lr_decay=0.8)
    t_tr = time.perf_counter() - t0
    del tr, T
    if args.save_weights:
        os.makedirs(os.path.dirname(os.path.abspath(args.save_weights)), exist_ok=True)
        np.savez_compressed(args.save_weights, weights=s.weights(), n_in=s.n_in, train_voxels=args.train_voxels,
                            epochs=args.epochs, loss=np.array(losses))
    nvox = args.nvox or synth.C4_NVOX
    v = synth.voxels("c4", nvox)
    N = v["bvals"].size
    x = F.Xfibres(**opts)
    x.set_acquisition(v["bvals"], v["bvecs"])
    x.load_data(v["data"])
    t = time.perf_counter()
    x.run(0)
    x.sync()
    plain_ms = (time.perf_counter() - t) * 1e3
    tp = x.timing()
    plain_out = x.outputs(raw_scalars=False)
    x.close()
    x = F.Xfibres(**opts)
    x.set_acquisition(v["bvals"], v["bvecs"])
    x.load_data(v["data"])
    t = time.perf_counter()
    nfb = x.run_surrogate(s, **SG_DEFAULTS_C4)
    learned_ms = (time.perf_counter() - t) * 1e3
    tl = x.timing()
    out = x.outputs(raw_scalars=False)
    x.close()
    s.close()
    e_l = SG.dyad_errors(out["dyads"], out["mean_f"], v["dirs"], v["f"])
    e_p = SG.dyad_errors(plain_out["dyads"], plain_out["mean_f"], v["dirs"], v["f"])
    # posterior f against the FP64 oracle (plain LM + full chains) on a random sample
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(nvox, args.oracle_voxels, replace=False))
    O.set_threads(min(16, os.cpu_count() or 1))
    oo = O.opts(**opts)
    t = time.perf_counter()
    d64 = v["data"][idx].astype(np.float64)
    p, tau, _ = O.fit(oo, v["bvals"], v["bvecs"], d64)
    smp = O.mcmc(oo, v["bvals"], v["bvecs"], d64, p, tau)
    fin = O.finish(oo, smp)
    t_orc = time.perf_counter() - t
    df_l = np.abs(np.sort(out["mean_f"][:, idx], 0) - np.sort(fin["mean_f"], 0))
    df_p = np.abs(np.sort(plain_out["mean_f"][:, idx], 0) - np.sort(fin["mean_f"], 0))
    fl_mcmc = flops_per_jump(3, 3, N, rician=True) * (opts["nburn"] + opts["njumps"]) * nvox
    line = {
        "metric": "voxels/s, learned start, model 3 + Rician, high-resolution volume", "value": nvox / (learned_ms / 1e3),
        "unit": "voxels/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": learned_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 chains, bf16 network",
        "data": "synthetic (seeded Philox), no published dataset",
        "config": {"workload": f"model 3 + --rician, nf=3, N={N}, {nvox} voxels (brain-shaped mask in 145x174x145, "
                               "SNR 10-30 smooth field, 15% of crossings < 45 deg), learned start burn-in 200",
                   "plain_voxels_per_s": nvox / (plain_ms / 1e3), "speedup_vs_plain": plain_ms / learned_ms,
                   "fallback_frac": nfb / nvox,
                   "dyad_err_median_deg": {"learned": float(np.nanmedian(e_l)), "plain": float(np.nanmedian(e_p))},
                   "posterior_f_vs_fp64_oracle": {"voxels": int(args.oracle_voxels), "oracle_s": t_orc,
                                                  "learned_abs_df_median": float(np.median(df_l)),
                                                  "learned_abs_df_p95": float(np.percentile(df_l, 95)),
                                                  "plain_abs_df_median": float(np.median(df_p)),
                                                  "plain_abs_df_p95": float(np.percentile(df_p, 95))},
                   "plain_ms": {"fit": tp["ms_fit"], "mcmc": tp["ms_mcmc"]},
                   "plain_mcmc_tflops": fl_mcmc / (tp["ms_mcmc"] / 1e3) / 1e12,
                   "plain_mcmc_frac_fp32_peak": fl_mcmc / (tp["ms_mcmc"] / 1e3) / 1e12 / FP32_PEAK_TFLOPS,
                   "learned_ms": {"infer": tl["ms_fit"], "chains_and_fallback": tl["ms_mcmc"]},
                   "training": {"voxels": args.train_voxels, "epochs": args.epochs, "final_loss": losses[-1],
                                "s_synth": t_syn, "s_label": t_lab, "s_train": t_tr}},
    }
    print(json.dumps(line), flush=True)


C2_MIX = [("c2_33", 1, 1), ("c2_69", 2, 2), ("c2_145", 2, 3), ("c2_288", 3, 3)]


def bench_gradnonlin(args):
    """The multi-shell volume acquired through per-voxel gradient deviations
    (grad_dev entries ~ N(0, 0.03)) and run with --gradnonlin: every voxel is
    fitted and sampled with its own corrected acquisition.  Also runs the same
    data with the shared acquisition (deviations ignored) for the cost and the
    accuracy the correction buys."""
    import fdt_amd as F
    from fdt_amd import synth
    from fdt_amd.surrogate import dyad_errors

    nvox = args.nvox or synth.C1_NVOX
    t0 = time.perf_counter()
    v = synth.gradnonlin_voxels("c1", nvox, sd=0.03)
    t_syn = time.perf_counter() - t0
    N = v["bvals"].size
    opts = dict(nfibres=3, model=2, nburn=1000, njumps=1250, sampleevery=25)

    def run(grad):
        x = F.Xfibres(**opts)
        try:
            x.set_acquisition(v["bvals"], v["bvecs"])
            x.load_data(v["data"])
            if grad is not None:
                x.load_grad_dev(grad)
            for _ in range(args.warmup):
                x.run(0)
                x.sync()
            ms = []
            t = time.perf_counter()
            for _ in range(args.steps):
                x.run(0)
                x.sync()
                ms.append(x.timing())
            step = (time.perf_counter() - t) / args.steps
            out = x.outputs(raw_scalars=False)
        finally:
            x.close()
        e = dyad_errors(out["dyads"], out["mean_f"], v["dirs"], v["f"])
        return dict(rate=nvox / step, ms_step=step * 1e3, ms_fit=float(np.mean([m["ms_fit"] for m in ms])),
                    ms_mcmc=float(np.mean([m["ms_mcmc"] for m in ms])), dyad=float(np.nanmedian(e)),
                    unmatched=float(np.mean(np.isnan(e))))

    g = run(v["grad"])
    p = run(None)
    fl_launch = flops_per_jump(2, 3, N) * (opts["nburn"] + opts["njumps"]) * nvox
    achieved = fl_launch / (g["ms_mcmc"] / 1e3) / 1e12
    line = {
        "metric": "voxels/s, 3-fibre multi-shell bedpostx with --gradnonlin", "value": g["rate"],
        "unit": "voxels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": g["ms_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded Philox), no published dataset",
        "config": {"workload": f"multi-shell model 2, nf=3, N={N}, {nvox} voxels acquired through per-voxel "
                               "gradient deviations (grad_dev ~ N(0, 0.03)), --gradnonlin, bi=1000 nj=1250 se=25",
                   "ms_fit": g["ms_fit"], "ms_mcmc": g["ms_mcmc"], "s_synth": t_syn},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "mcmc_kernel<3,18,false,true>",
                     "flops_per_launch": fl_launch},
        "shared_acquisition": {"voxels_per_s": p["rate"], "ms_fit": p["ms_fit"], "ms_mcmc": p["ms_mcmc"]},
        "dyad_err_median_deg": {"gradnonlin": g["dyad"], "deviations_ignored": p["dyad"]},
        "unmatched_fibres": {"gradnonlin": g["unmatched"], "deviations_ignored": p["unmatched"]},
    }
    print(json.dumps(line), flush=True)


def bench_spatial(args):
    """The CPU xfibres initialisation (cpu_init): single-shell model 1, nf=2
    voxels on a slab of the brain-shaped 145x174x145 mask, run with neighbour
    initialisation (each chain started from the best finished neighbour, in
    wavefronts x + 2y + 4z), with the tensor start (--nospat), and with the
    xfibres_gpu constrained nonlinear fit on the same voxels."""
    import fdt_amd as F
    from fdt_amd import synth
    from fdt_amd.surrogate import dyad_errors

    nx, ny, nz = synth.C4_GRID
    pos = synth.c4_positions().astype(np.int64)
    zc = nz // 2
    sel = pos[np.abs(pos[:, 2] - zc) < args.slab / 2.0]
    gidx = np.sort(sel[:, 0] + nx * (sel[:, 1] + ny * sel[:, 2]))
    if args.nvox:
        gidx = gidx[:args.nvox]
    nvox = len(gidx)
    v = synth.voxels("c0", nvox)
    N = v["bvals"].size
    x_, y_, z_ = gidx % nx, (gidx // nx) % ny, gidx // (nx * ny)
    nwave = len(np.unique(x_ + 2 * y_ + 4 * z_))
    opts = dict(nfibres=2, model=1, nburn=1000, njumps=1250, sampleevery=25)

    def run(**kw):
        x = F.Xfibres(**opts, **kw)
        try:
            x.set_acquisition(v["bvals"], v["bvecs"])
            x.load_data(v["data"])
            x.set_grid((nx, ny, nz), gidx)
            for _ in range(args.warmup):
                x.run(0)
                x.sync()
            ms = []
            t = time.perf_counter()
            for _ in range(args.steps):
                x.run(0)
                x.sync()
                ms.append(x.timing())
            step = (time.perf_counter() - t) / args.steps
            out = x.outputs(raw_scalars=False)
        finally:
            x.close()
        e = dyad_errors(out["dyads"], out["mean_f"], v["dirs"], v["f"])
        return dict(rate=nvox / step, ms_step=step * 1e3, ms_fit=float(np.mean([m["ms_fit"] for m in ms])),
                    ms_mcmc=float(np.mean([m["ms_mcmc"] for m in ms])), dyad=float(np.nanmedian(e)),
                    unmatched=float(np.mean(np.isnan(e))))

    sp = run(cpu_init=1)
    ts = run(cpu_init=1, nospat=1)
    nl = run(cnonlin=1)
    conc = None
    if args.concurrent > 1:
        # several volumes (further draws of the same distribution on the same grid) at once, one context and
        # stream each: a neighbour-initialised volume leaves most CUs idle, so volumes share the GPU
        xs = []
        try:
            for j in range(args.concurrent):
                vj = synth.voxels("c0", nvox, sub=j + 1)
                x = F.Xfibres(**opts, cpu_init=1)
                xs.append(x)
                x.set_acquisition(vj["bvals"], vj["bvecs"])
                x.load_data(vj["data"])
                x.set_grid((nx, ny, nz), gidx)
            for x in xs:
                x.run(0)
            for x in xs:
                x.sync()
            t = time.perf_counter()
            for x in xs:
                x.run(0)
            for x in xs:
                x.sync()
            dt = time.perf_counter() - t
        finally:
            for x in xs:
                x.close()
        conc = {"volumes": args.concurrent, "voxels_per_s": args.concurrent * nvox / dt, "s": dt}
    line = {
        "metric": "voxels/s, 2-fibre single-shell xfibres with the CPU xfibres neighbour initialisation",
        "value": sp["rate"], "unit": "voxels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sp["ms_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded Philox), no published dataset",
        "config": {"workload": f"single-shell model 1, nf=2, N={N}, {nvox} voxels: {args.slab} central slices "
                               f"of the brain-shaped {nx}x{ny}x{nz} mask, bi=1000 nj=1250 se=25, cpu_init "
                               "(neighbour initialisation)",
                   "wavefronts": nwave, "voxels_per_wavefront": nvox / nwave,
                   "ms_per_wavefront": sp["ms_mcmc"] / nwave, "ms_chains": sp["ms_mcmc"]},
        "tensor_start": {"voxels_per_s": ts["rate"], "ms_start": ts["ms_fit"], "ms_mcmc": ts["ms_mcmc"]},
        "nonlinear_fit": {"voxels_per_s": nl["rate"], "ms_fit": nl["ms_fit"], "ms_mcmc": nl["ms_mcmc"]},
        "dyad_err_median_deg": {"neighbour": sp["dyad"], "tensor": ts["dyad"], "nonlinear": nl["dyad"]},
        "unmatched_fibres": {"neighbour": sp["unmatched"], "tensor": ts["unmatched"], "nonlinear": nl["unmatched"]},
    }
    if conc:
        line["concurrent_volumes"] = conc
    print(json.dumps(line), flush=True)


def _write_nii(path, arr, pixdim=(2.0, 2.0, 2.0)):
    """uncompressed single-file NIfTI-1, float32 or uint8, arr indexed [x, y, z(, t)]"""
    import struct
    arr = np.asarray(arr)
    code = {np.dtype(np.float32): 16, np.dtype(np.uint8): 2}[arr.dtype]
    h = bytearray(352)
    struct.pack_into("<i", h, 0, 348)
    struct.pack_into("<8h", h, 40, *([arr.ndim] + list(arr.shape) + [1] * (7 - arr.ndim)))
    struct.pack_into("<2h", h, 70, code, arr.dtype.itemsize * 8)
    struct.pack_into("<8f", h, 76, 1.0, *pixdim, 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<f", h, 108, 352.0)
    h[344:348] = b"n+1\0"
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(np.asfortranarray(arr).tobytes(order="F"))


def bench_pipeline(args):
    """The whole bedpostx_gpu pipeline from files: the multi-shell volume as a
    NIfTI subject (110 x 110 x 72 grid, ellipsoidal mask of 250,968 voxels,
    N = 288) through fdt_amd/bin/bedpostx_gpu with the bedpostx defaults
    (nf 3, model 2, bi 1000, nj 1250, se 25) to the .bedpostX images."""
    import shutil
    import subprocess
    import tempfile
    from fdt_amd import synth

    grid, axes = (110, 110, 72), (48, 48, 26)
    c = [(n - 1) / 2.0 for n in grid]
    x, y, z = np.meshgrid(*[np.arange(n) for n in grid], indexing="ij")
    mask = (((x - c[0]) / axes[0]) ** 2 + ((y - c[1]) / axes[1]) ** 2 + ((z - c[2]) / axes[2]) ** 2) <= 1.0
    nvox = int(mask.sum())
    v = synth.voxels("c1", nvox, with_truth=False)
    N = v["bvals"].size
    xs, ys, zs = np.nonzero(mask.transpose(2, 1, 0))[::-1]
    root = tempfile.mkdtemp(prefix="fdt_pipeline_")
    try:
        subj = os.path.join(root, "subj")
        os.makedirs(subj)
        t0 = time.perf_counter()
        data = np.zeros(grid + (N,), np.float32)
        data[xs, ys, zs, :] = v["data"]
        _write_nii(os.path.join(subj, "data.nii"), data)
        del data
        _write_nii(os.path.join(subj, "nodif_brain_mask.nii"), mask.astype(np.uint8))
        np.savetxt(os.path.join(subj, "bvals"), v["bvals"][None, :])
        np.savetxt(os.path.join(subj, "bvecs"), v["bvecs"])
        t_prep = time.perf_counter() - t0
        pipe = os.path.join(HERE, "fdt_amd", "bin", "bedpostx_gpu")
        t0 = time.perf_counter()
        r = subprocess.run([pipe, subj], capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"bedpostx_gpu failed: {r.stderr[-2000:]}")
        log = open(os.path.join(root, "subj.bedpostX", "logs", "bedpostx_gpu.log")).read()
        out_dir = os.path.join(root, "subj.bedpostX")
        out_bytes = sum(os.path.getsize(os.path.join(out_dir, f)) for f in os.listdir(out_dir)
                        if f.endswith(".nii.gz"))
        n_img = sum(1 for f in os.listdir(out_dir) if f.endswith(".nii.gz"))
    finally:
        shutil.rmtree(root, ignore_errors=True)

    def field(prefix):
        for line in log.splitlines():
            if line.startswith(prefix):
                return float(line.split(":")[1].split()[0])
        return None
    line = {
        "metric": "voxels/s, 3-fibre multi-shell bedpostx_gpu pipeline, NIfTI files in to .bedpostX images out",
        "value": nvox / wall, "unit": "voxels/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded Philox), no published dataset",
        "config": {"workload": f"bedpostx_gpu <subject> with defaults (nf 3, model 2, bi 1000, nj 1250, se 25): "
                               f"110x110x72 grid, {nvox} mask voxels, N={N}, uncompressed data.nii in, "
                               f"{n_img} .nii.gz images out ({out_bytes / 1e6:.1f} MB)",
                   "s_read_and_mask": field("read + mask"), "s_engine": field("engine"),
                   "s_write": field("write"), "s_total_in_process": field("total"), "s_wall": wall,
                   "s_make_subject": t_prep},
    }
    print(json.dumps(line), flush=True)


def bench_c2(args):
    """Mixed queue of four volumes (N = 33/69/145/288, models 1/2/2/3, nf
    1/2/3/3, 60,000 voxels each): the batched path (concurrent LM, one
    persistent chain launch, hipGraph replay, autotuned tiles) against the
    same volumes run one after the other through the per-volume path."""
    import fdt_amd as F
    from fdt_amd import synth
    nvox = args.nvox or 60000
    run = dict(nburn=1000, njumps=1250, sampleevery=25)
    vols = [(synth.voxels(name, nvox, with_truth=False), dict(run, nfibres=K, model=m)) for name, m, K in C2_MIX]
    # separate: one volume at a time
    xs = []
    for j, (v, kw) in enumerate(vols):
        x = F.Xfibres(**kw)
        x.set_acquisition(v["bvals"], v["bvecs"])
        x.load_data(v["data"])
        xs.append(x)
    for x in xs:  # warm-up
        x.run(0)
        x.sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        for j, x in enumerate(xs):
            x.run(j * nvox)
            x.sync()
    sep = (time.perf_counter() - t) / args.steps
    for x in xs:
        x.close()
    # batched
    q = F.Queue()
    for j, (v, kw) in enumerate(vols):
        q.enqueue(v["bvals"], v["bvecs"], v["data"], vox_offset=j * nvox, **kw)
    tune_ms = q.autotune(jumps=50, sample_vox=4096)
    q.run()  # eager run + graph capture
    for _ in range(args.warmup):
        q.run()
    t = time.perf_counter()
    for _ in range(args.steps):
        q.run()
    bat = (time.perf_counter() - t) / args.steps
    tm = q.timing()
    q.close()
    total = nvox * len(C2_MIX)
    line = {
        "metric": "voxels/s, mixed-acquisition queue (batched vs one volume at a time)", "value": total / bat,
        "unit": "voxels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": bat * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded Philox), no published dataset",
        "config": {"workload": f"4 volumes x {nvox} voxels: (N, model, nf) = (33,1,1) (69,2,2) (145,2,3) (288,3,3), "
                               "bi=1000 nj=1250 se=25",
                   "separate_ms": sep * 1e3, "speedup_vs_separate": sep / bat,
                   "autotuned": {"rows_per_wg": tm["rows_per_wg"], "vpl": tm["vpl"], "tune_ms": tune_ms},
                   "graph": "hipGraph replay of LM stages + queue kernel + finishing"},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nvox", type=int, default=0, help="override the voxel count (default: the full mask)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="time the plain path only (no end-to-end pass, no learned start): for profiler runs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", default="c1", choices=["c1", "c2", "c3", "c4", "gradnonlin", "pipeline", "spatial"],
                    help="c1: the multi-shell 3-fibre volume (default); c2: the mixed-acquisition queue; "
                         "c3: train the learned start and compare it with the plain path; c4: the same for "
                         "model 3 + Rician on the 904,550-voxel volume; gradnonlin: the multi-shell volume "
                         "with per-voxel gradient deviations; pipeline: the bedpostx_gpu pipeline from NIfTI "
                         "files to the .bedpostX images; spatial: the CPU xfibres neighbour initialisation "
                         "on a slab of the brain-shaped mask")
    ap.add_argument("--slab", type=int, default=16, help="spatial: central z slices of the mask")
    ap.add_argument("--concurrent", type=int, default=1, help="spatial: also run this many volumes at once")
    ap.add_argument("--oracle-voxels", type=int, default=64)
    ap.add_argument("--train-voxels", type=int, default=2097152)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--save-weights", default="")
    ap.add_argument("--lm-mode", type=int, default=LM_DEFAULT, choices=[0, 1, 2],
                    help="LM arithmetic: 0 FP32, 1 FP64, 2 FP32 normal equations with FP64 parameters and cost")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this engine runs on one GPU per process "
              "(voxels shard by part across processes as bedpostx_gpu does); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c2":
        bench_c2(args)
        return
    if args.workload == "c3":
        bench_c3(args)
        return
    if args.workload == "c4":
        bench_c4(args)
        return
    if args.workload == "gradnonlin":
        bench_gradnonlin(args)
        return
    if args.workload == "pipeline":
        bench_pipeline(args)
        return
    if args.workload == "spatial":
        bench_spatial(args)
        return

    import fdt_amd as F
    from fdt_amd import synth

    nvox = args.nvox or synth.C1_NVOX
    v = synth.voxels("c1", nvox)
    N = v["bvals"].size
    opts = dict(nfibres=3, model=2, nburn=1000, njumps=1250, sampleevery=25, lm_fp64=args.lm_mode)
    x = F.Xfibres(**opts)
    x.set_acquisition(v["bvals"], v["bvecs"])
    x.load_data(v["data"])

    for _ in range(args.warmup):
        x.run(0)
        x.sync()
    ms_mcmc, ms_fit, ms_fin = [], [], []
    x.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x.run(0)
        x.sync()
        t = x.timing()
        ms_mcmc.append(t["ms_mcmc"]); ms_fit.append(t["ms_fit"]); ms_fin.append(t["ms_finish"])
    x.sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    value = nvox / (ms_step / 1e3)
    tm = x.timing()
    iters = x.lm_iters()
    plain_out = x.outputs(raw_scalars=False)
    x.close()

    # the same steps from host buffers: upload of the signal, LM, chains,
    # finishing, download of every output (the PCIe-inclusive rate), with the
    # signal and the outputs in page-locked host arrays allocated once
    e2e_ms = float("nan")
    if not args.plain_only:
        x2 = F.Xfibres(**opts)
        x2.set_acquisition(v["bvals"], v["bvecs"])
        host_data = F.pinned_empty(v["data"].shape)
        host_data[...] = v["data"]
        x2.load_data(host_data)
        host_out = x2.output_buffers(alloc=F.pinned_empty)
        x2.sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            x2.load_data(host_data)
            x2.run(0)
            x2.outputs(into=host_out)
        e2e_ms = (time.perf_counter() - t) / args.steps * 1e3
        x2.close()
        del host_data, host_out

    fl_launch = flops_per_jump(2, 3, N) * (opts["nburn"] + opts["njumps"]) * nvox
    mc_ms = float(np.mean(ms_mcmc))
    achieved = fl_launch / (mc_ms / 1e3) / 1e12
    traffic = MCMC_TRAFFIC_BYTES if nvox == synth.C1_NVOX else None
    lm_ms = float(np.mean(ms_fit))
    lm_flops = lm_flops_per_iter(N, 12) * float(iters.sum())
    hist_edges = [0, 20, 40, 60, 80, 100, 150, 200, 300, 400, 10 ** 6]
    hist = np.histogram(iters, bins=hist_edges)[0]
    line = {
        "metric": METRIC, "value": value, "unit": "voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded Philox), no published dataset",
        "value_scope": "signal resident in HBM when the step starts; LM + 2,250 chain sweeps + finishing per "
                       "step; PCIe transfers excluded (see end_to_end)",
        "config": {"workload": f"multi-shell model 2, nf=3, N={N} (18 b0 + 3x90 at b=1000/2000/3000), "
                               f"{nvox} voxels, bi=1000 nj=1250 se=25, plain LM init",
                   "init": "plain LM (tensor start -> PVM_single -> PVM_multi)",
                   "lm_precision": LM_MODES[args.lm_mode],
                   "ms_fit": lm_ms, "ms_mcmc": mc_ms, "ms_finish": float(np.mean(ms_fin)),
                   "mcmc_tile": f"{tm['lanes']} lanes x {tm['vpl']} volumes per voxel"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "mcmc_kernel<3,18,false,false>", "flops_per_launch": fl_launch,
                     "compute": "FP32 vector ALU (no matrix-shaped work on this path); peak is the published "
                                "FP32 figure, which counts packed FP32 -- this build disables packed FP32, so "
                                "half of it (78.6) is the reachable issue ceiling"},
        "end_to_end": {"voxels_per_s": nvox / (e2e_ms / 1e3), "ms_per_step": e2e_ms, "steps": args.steps,
                       "scope": "from page-locked host buffers (allocated once): upload of the 288-volume "
                                "signal, LM, chains, finishing, download of every sample and mean"},
        "lm": {"ms": lm_ms, "iters_mean": float(iters.mean()), "iters_median": float(np.median(iters)),
               "iters_hist": {f"{a}-{b - 1}" if b < 10 ** 6 else f">={a}": int(c)
                              for a, b, c in zip(hist_edges[:-1], hist_edges[1:], hist)},
               "achieved_tflops": lm_flops / (lm_ms / 1e3) / 1e12,
               "flop_model": "N*(np(np+1) + 2np + 40K) FP32 per loop trip at np=12, K=3 (both stages counted at "
                             "the stage-2 size)"},
    }
    if os.path.exists(SURROGATE_WEIGHTS) and not args.plain_only:
        import fdt_amd as F2
        wz = np.load(SURROGATE_WEIGHTS)
        s = F2.Surrogate(N, seed=0)
        s.load(wz["weights"])
        sg = surrogate_eval(v, opts, s, args.steps, plain_out, SG_DEFAULTS)
        sg50 = surrogate_eval(v, opts, s, 1, None, SG_SHORT)
        s.close()
        sg["speedup_vs_plain"] = sg["value"] / value
        sg50["speedup_vs_plain"] = sg50["value"] / value
        sg["burnin50"] = sg50
        sg["ablation_plain_lm_short_burnin"] = {str(bi): plain_short_burnin(v, opts, bi) for bi in (200, 50)}
        rec = opts["njumps"] / (opts["nburn"] + opts["njumps"])
        floor_ms = mc_ms * rec + float(np.mean(ms_fin))
        sg["speedup_bound"] = (f"with the {opts['njumps']:,} recorded sweeps kept, a pass costs at least their chain "
                               f"time ({rec:.2f} of the plain chain launch) plus finishing, {floor_ms:.0f} ms, so no "
                               f"start can exceed {ms_step / floor_ms:.2f}x over the plain path here")
        sg["weights"] = (f"fdt_amd/weights/{os.path.basename(SURROGATE_WEIGHTS)}: trained by bench.py --workload c3 "
                         f"on {int(wz['train_voxels'])} voxels, {int(wz['epochs'])} epochs")
        line["surrogate"] = sg
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(v, opts, seconds=args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
