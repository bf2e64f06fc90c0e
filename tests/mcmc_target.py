"""Closed-form targets of the xfibres MCMC (fibre.h:283-320, :931-953;
CUDA/runmcmc_kernels.cu:118-191, :332-375).

With every volume at b = 0 the signal model is S0 for every parameter value
(all isotropic and stick signals are exactly 1), so the likelihood depends on
S0 alone and every other sampled parameter follows its prior inside its
bounds.  With proposal-width adaptation switched off (update interval longer
than the chain) the Metropolis chains then sample these closed forms:

  |cos theta| ~ U(0, 1)          theta prior -log|sin theta / 2|
  d ~ U(0, 0.005)                models 1/2: flat prior, bounds [0, 0.005]
  d ~ Gamma(3, 1/4000) | d<=0.005  model 3: prior -2 log d + 4000 d
  R ~ N(Rmean, Rstd/sqrt 2) | (Rmean-2 Rstd, 2 Rmean]   model 3 R prior
  f ~ Beta(1/2, 1)               K=1, --allard --fudge 0.5: prior 0.5 log f
  S0 ~ ybar + sqrt(SS_y / (N (N-1))) t_{N-1}   Gaussian energy (N/2) log SS

Each check also rejects the distribution a plausible mistake would produce
(a dropped theta prior, a flipped ARD sign, the likelihood exponent N
instead of N/2), so the tests have the power to see those mistakes.  Shared
by the oracle tests (CPU) and the HIP chain tests (GPU)."""
from __future__ import annotations

import numpy as np
from scipy import stats

N = 30
UPPERDIFF = 0.005
NO_ADAPT = 1_000_000_000   # proposal update interval longer than any chain here
RUN = dict(nburn=2000, njumps=4000, sampleevery=1000)


def acquisition():
    bvals = np.zeros(N)
    bvecs = np.zeros((3, N))
    bvecs[0] = 1.0
    return bvals, bvecs


def data(nvox: int, seed: int = 5):
    g = np.random.default_rng(seed)
    S0 = g.uniform(800, 1200, nvox)
    return S0[:, None] + g.standard_normal((nvox, N)) * 40.0


def start(model: int, K: int, y: np.ndarray, f=0.3, d=None):
    """Chain start vectors [S0, d, (d_std), (f, th, ph) x K] with the
    parameter layout of the LM output."""
    nvox = y.shape[0]
    add = 1 if model >= 2 else 0
    p = np.zeros((nvox, 2 + add + 3 * K))
    p[:, 0] = y.mean(1)
    p[:, 1] = (7.5e-4 if model == 3 else 2.5e-3) if d is None else d
    if add:
        p[:, 2] = 5e-4
    g = np.random.default_rng(11)
    for k in range(K):
        p[:, 2 + add + 3 * k] = f / K
        p[:, 3 + add + 3 * k] = g.uniform(0.2, 2.9, nvox)
        p[:, 4 + add + 3 * k] = g.uniform(-3.0, 3.0, nvox)
    return p


def opts(model: int, K: int, **kw):
    o = dict(nfibres=K, model=model, updateproposalevery=NO_ADAPT, **RUN)
    o.update(kw)
    return o


def ks(x, cdf):
    return stats.kstest(np.ravel(x), cdf).pvalue


def check_theta(th):
    """|cos theta| is uniform; a flat theta prior would give the arcsine law."""
    c = np.abs(np.cos(np.asarray(th, np.float64)))
    good = ks(c, stats.uniform.cdf)
    flat = ks(c, lambda x: 2 / np.pi * np.arcsin(np.clip(x, 0, 1)))
    return good, flat


def check_d_uniform(d):
    good = ks(d, stats.uniform(0, UPPERDIFF).cdf)
    # a Gamma(3, 1/4000) prior applied by mistake
    wrong = ks(d, lambda x: stats.gamma(3, scale=1 / 4000).cdf(x) / stats.gamma(3, scale=1 / 4000).cdf(UPPERDIFF))
    return good, wrong


def check_d_gamma(d):
    G = stats.gamma(3, scale=1 / 4000)
    good = ks(d, lambda x: G.cdf(x) / G.cdf(UPPERDIFF))
    # the 4000 d term dropped: density d^2 on (0, 0.005]
    wrong = ks(d, lambda x: (np.clip(x, 0, UPPERDIFF) / UPPERDIFF) ** 3)
    return good, wrong


def check_R(R, Rmean=0.13, Rstd=0.03):
    lo, hi = Rmean - 2 * Rstd, 2 * Rmean
    Nm = stats.norm(Rmean, Rstd / np.sqrt(2))
    good = ks(R, lambda x: (Nm.cdf(np.clip(x, lo, hi)) - Nm.cdf(lo)) / (Nm.cdf(hi) - Nm.cdf(lo)))
    # the prior read as a normal of width Rstd
    Nw = stats.norm(Rmean, Rstd)
    wrong = ks(R, lambda x: (Nw.cdf(np.clip(x, lo, hi)) - Nw.cdf(lo)) / (Nw.cdf(hi) - Nw.cdf(lo)))
    return good, wrong


def check_f_beta(f):
    good = ks(f, stats.beta(0.5, 1).cdf)
    flipped = ks(f, stats.beta(1.5, 1).cdf)   # ARD prior with the opposite sign
    return good, flipped


def check_S0(S0, y):
    """Standardised S0 draws against Student t with N-1 degrees of freedom."""
    S0 = np.asarray(S0, np.float64)           # [ns][nvox]
    ybar = y.mean(1)
    ss = ((y - ybar[:, None]) ** 2).sum(1)
    z = (S0 - ybar[None, :]) / np.sqrt(ss / (N * (N - 1)))[None, :]
    good = ks(z, stats.t(N - 1).cdf)
    # energy N log SS (exponent doubled): S0 ~ ybar + sqrt(SS/(N(2N-1))) t_{2N-1}
    wrong = ks(z, stats.t(2 * N - 1, scale=np.sqrt((N - 1) / (2 * N - 1))).cdf)
    return good, wrong
