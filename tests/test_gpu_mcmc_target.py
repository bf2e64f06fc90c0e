"""The HIP chain samples the closed-form targets of tests/mcmc_target.py
(all-b=0 acquisition, fixed proposal widths): theta, d, R and f follow their
priors and S0 its Student-t posterior, as the oracle's chain does
(tests/test_oracle_mcmc_target.py).  N = 30 also leaves padding in the middle
slot of the 3-slot tile."""
import numpy as np
import pytest

import mcmc_target as T

pytestmark = pytest.mark.gpu

NVOX = 2048
P_OK, P_BAD = 1e-3, 1e-6


def run(fdt, model, K, rician=0, **kw):
    y = T.data(NVOX)
    bvals, bvecs = T.acquisition()
    x = fdt.Xfibres(**T.opts(model, K, **kw))
    try:
        x.set_acquisition(bvals, bvecs)
        x.load_data(y.astype(np.float32))
        x.set_params(T.start(model, K, y).astype(np.float32), None)
        x.mcmc(0)
        x.finish()
        out = x.outputs()
    finally:
        x.close()
    return y.astype(np.float32).astype(np.float64), out


def test_theta_prior_and_model1_d_and_S0(fdt):
    y, s = run(fdt, 1, 1)
    good, flat = T.check_theta(s["th"])
    assert good > P_OK and flat < P_BAD, (good, flat)
    good, wrong = T.check_d_uniform(s["d"])
    assert good > P_OK and wrong < P_BAD, (good, wrong)
    good, wrong = T.check_S0(s["S0"], y)
    assert good > P_OK and wrong < P_BAD, (good, wrong)


def test_ard_prior_beta(fdt):
    y, s = run(fdt, 1, 1, all_ard=1, fudge=0.5)
    good, flipped = T.check_f_beta(s["f"])
    assert good > P_OK and flipped < P_BAD, (good, flipped)


def test_model3_d_gamma_and_R_prior(fdt):
    y, s = run(fdt, 3, 1)
    good, wrong = T.check_d_gamma(s["d"])
    assert good > P_OK and wrong < P_BAD, (good, wrong)
    good, wrong = T.check_R(s["R"])
    assert good > P_OK and wrong < P_BAD, (good, wrong)


def test_three_fibres_model2(fdt):
    """K=3: every fibre's theta follows its prior (the f of fibres 2 and 3
    are under an improper ARD prior and are not checked)."""
    y, s = run(fdt, 2, 3)
    for k in range(3):
        good, flat = T.check_theta(s["th"][k])
        assert good > P_OK and flat < P_BAD, (k, good, flat)
    good, wrong = T.check_S0(s["S0"], y)
    assert good > P_OK and wrong < P_BAD, (good, wrong)
