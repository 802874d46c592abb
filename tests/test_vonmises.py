"""The seeded link draws every BASELINE link field comes from (qed3_random_gauge, host): von
Mises(0, kappa) by Best & Fisher's rejection method, uniform on [-pi, pi) for kappa <= 0.
Checked against the distribution itself: E[cos theta] = I1(kappa) / I0(kappa), E[sin theta]
= 0, and a Kolmogorov-Smirnov test against scipy's von Mises CDF."""
import numpy as np
import pytest
import scipy.special as sp
import scipy.stats as st

import qed3_amd as q


@pytest.mark.parametrize("kappa", [0.5, 1.4, 2.0, 6.0])
def test_von_mises_moments_and_distribution(kappa):
    sp_, tp_ = q.random_gauge(0, 16, 16, seed=int(10 * kappa), kappa=kappa)
    th = np.concatenate([sp_.ravel(), tp_.ravel()])
    assert th.min() >= -np.pi and th.max() <= np.pi
    n = th.size
    m1 = sp.i1(kappa) / sp.i0(kappa)
    # standard error of the mean of cos: sqrt(Var cos / n), Var cos = (1 + I2/I0)/2 - m1^2
    var = 0.5 * (1 + sp.iv(2, kappa) / sp.i0(kappa)) - m1 ** 2
    assert abs(np.cos(th).mean() - m1) < 5 * np.sqrt(var / n)
    assert abs(np.sin(th).mean()) < 5 * np.sqrt((1 - (var + m1 ** 2)) / n)
    ks = st.kstest(th, st.vonmises(kappa).cdf)
    assert ks.pvalue > 1e-4, ks


def test_uniform_links():
    sp_, tp_ = q.random_gauge(1, 8, 8, seed=3, kappa=0.0)
    th = np.concatenate([sp_.ravel(), tp_.ravel()])
    assert th.min() >= -np.pi and th.max() < np.pi
    ks = st.kstest(th, st.uniform(-np.pi, 2 * np.pi).cdf)
    assert ks.pvalue > 1e-4, ks


def test_a_wrong_concentration_is_detected():
    """The moment check separates neighbouring kappas at this sample size."""
    sp_, tp_ = q.random_gauge(0, 16, 16, seed=14, kappa=1.4)
    th = np.concatenate([sp_.ravel(), tp_.ravel()])
    assert abs(np.cos(th).mean() - sp.i1(1.3) / sp.i0(1.3)) > 0.01
