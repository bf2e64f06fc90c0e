"""The oracle's MCMC target pinned by closed forms (tests/mcmc_target.py):
with an all-b=0 acquisition and fixed proposal widths the chains sample the
priors of theta, d, R and f in closed form and a Student-t posterior of S0,
which pins the theta prior (fibre.h:283-290), the ARD prior and its sign
(fibre.h:300-320), the model-3 d and R priors (fibre.h:816-843) and the
Gaussian energy (N/2) log SS (fibre.h:931-953).  A behavioural ARD check
(single-fibre voxels fitted with two fibres) completes the pin of the ARD
switch."""
import numpy as np
import pytest

from fdt_amd import synth
import mcmc_target as T

NVOX = 512
P_OK, P_BAD = 1e-3, 1e-6


def run(orc, model, K, nvox=NVOX, **kw):
    y = T.data(nvox)
    bvals, bvecs = T.acquisition()
    o = orc.opts(**T.opts(model, K, **kw))
    p = T.start(model, K, y)
    s = orc.mcmc(o, bvals, bvecs, y, p, None)
    return y, s


def test_theta_prior_and_model1_d_and_S0(orc):
    y, s = run(orc, 1, 1)
    good, flat = T.check_theta(s["th"])
    assert good > P_OK and flat < P_BAD, (good, flat)
    good, wrong = T.check_d_uniform(s["d"])
    assert good > P_OK and wrong < P_BAD, (good, wrong)
    good, wrong = T.check_S0(s["S0"], y)
    assert good > P_OK and wrong < P_BAD, (good, wrong)


def test_ard_prior_beta(orc):
    y, s = run(orc, 1, 1, all_ard=1, fudge=0.5)
    good, flipped = T.check_f_beta(s["f"])
    assert good > P_OK and flipped < P_BAD, (good, flipped)


def test_model3_d_gamma_and_R_prior(orc):
    y, s = run(orc, 3, 1)
    good, wrong = T.check_d_gamma(s["d"])
    assert good > P_OK and wrong < P_BAD, (good, wrong)
    good, wrong = T.check_R(s["R"])
    assert good > P_OK and wrong < P_BAD, (good, wrong)
    good, flat = T.check_theta(s["th"])
    assert good > P_OK and flat < P_BAD, (good, flat)


def single_fibre_voxels(nvox, seed=7):
    """Model-1 voxels with one stick (f ~ U(0.4, 0.7)) on the single-shell
    acquisition, Rician SNR 20."""
    v = synth.voxels("c0", 1)
    bvals, bvecs = v["bvals"], v["bvecs"]
    g = np.random.default_rng(seed)
    S0 = g.uniform(800, 1200, nvox)
    d = g.uniform(1e-3, 2e-3, nvox)
    f = np.zeros((nvox, 1))
    f[:, 0] = g.uniform(0.4, 0.7, nvox)
    dirs = g.standard_normal((nvox, 1, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    clean = synth.clean_signal(1, bvals, bvecs, S0, d, None, f, dirs)
    sig = S0[:, None] / 20
    y = np.hypot(clean + g.standard_normal(clean.shape) * sig, g.standard_normal(clean.shape) * sig)
    return bvals, bvecs, y


@pytest.mark.slow
def test_ard_switch_behaviour(orc):
    """One-fibre truth fitted with --nf=2: with ARD the second fraction is
    driven to ~0; without ARD it is not."""
    bvals, bvecs, y = single_fibre_voxels(96)
    frac = {}
    for noard in (0, 1):
        o = orc.opts(nfibres=2, model=1, no_ard=noard, nburn=1000, njumps=1250, sampleevery=25)
        p, tau, _ = orc.fit(o, bvals, bvecs, y)
        s = orc.mcmc(o, bvals, bvecs, y, p, tau)
        fin = orc.finish(o, s)
        frac[noard] = (fin["mean_f"][1] < 0.05).mean()
    print(f"mean f2 < 0.05: ARD {frac[0]:.3f}, no ARD {frac[1]:.3f}")
    assert frac[0] >= 0.9
    assert frac[1] <= frac[0] - 0.3
