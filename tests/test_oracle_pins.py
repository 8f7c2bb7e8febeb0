"""Pins of the oracle's denoiser-side functions by what the mathematics
fixes, not by their own formulas.

* Bonett's log-variance (lnS2, se2): on normal samples, lnS2 = ln(c s^2) has
  the exact mean ln(c sigma^2) + psi((n-1)/2) - ln((n-1)/2) (s^2 is a scaled
  chi-square), se2 (n-1)/c^2 = b2 - (n-3)/n has the exact mean
  3(n-1)/(n+1) - (n-3)/n (E[b2] for normal samples), and the equality test
  between two samples of one normal population rejects at about its nominal
  level 2(1 - Phi(2.80703)) = 0.005 (Bonett's test is asymptotic: the
  simulated rate is 0.0047 at n = 16 and 0.0058 at n = 64).  Writing c for
  c^2, n for n - 1, or K - 3 for K - (n-3)/n moves one of these far outside
  its tolerance.
* The range and spatial weights exp(ds |dp|^2 + sum_k dr_k |g_p - g_q|^2)
  (estimator.h:353, estimator.cpp:10-18): films of 1x3 and 3x3 pixels whose
  features differ by whole multiples of the feature's sd, with a nonzero
  spatial factor and every test accepting, give outputs computed here by
  hand from the exponents.
* Mode 3 (FILM_MULTI_TRANSFORM_MODE = 3): only the set with the smallest jb
  of the centre pixel is tested; a two-pixel film whose sets disagree gives
  the known answers both ways round, and mode 2 rejects.
"""
import numpy as np
import pytest
from scipy.special import digamma, polygamma

import oracle

Z = 2.80703


def _prepass_of_samples(x):
    """oracle.prepass of one set from samples x [P, n] (every channel the same)."""
    P, n = x.shape
    mu = x.mean(1)
    d = x - mu[:, None]
    st = np.zeros((1, 7, 3, 1, P), np.float32)
    for k, v in ((0, mu), (1, (d ** 2).sum(1)), (2, (d ** 3).sum(1)), (3, (d ** 4).sum(1))):
        st[0, k] = v.astype(np.float32)[None, None, :]
    jb, lnS2, se2, m2C = oracle.prepass(np.full((1, P), n, np.int32), st, Z)
    return lnS2[0, 0, 0], se2[0, 0, 0]


_SIM = {}


def _simulate(n, pairs, sigma=3.0, seed=0):
    if (n, pairs) not in _SIM:
        rng = np.random.default_rng(1000 + n)
        ls, ss = [], []
        for k in range(0, 2 * pairs, 40000):
            m = min(40000, 2 * pairs - k)
            l, s = _prepass_of_samples(sigma * rng.standard_normal((m, n)))
            ls.append(l)
            ss.append(s)
        _SIM[(n, pairs)] = (np.concatenate(ls), np.concatenate(ss))
    return _SIM[(n, pairs)]


CASES = [(16, 100000), (64, 50000), (256, 20000), (1024, 10000)]


@pytest.mark.parametrize("n,pairs", CASES)
def test_bonett_lnS2_and_se2_have_their_normal_theory_means(n, pairs):
    sigma = 3.0
    lnS2, se2 = _simulate(n, pairs, sigma)
    k = (n - 1) / 2.0
    c = n / (n - Z)
    want_l = np.log(c * sigma ** 2) + digamma(k) - np.log(k)
    sd_l = np.sqrt(polygamma(1, k) / lnS2.size)
    assert abs(lnS2.mean() - want_l) < 5 * sd_l, (lnS2.mean() - want_l, sd_l)
    q = se2 * (n - 1) / c ** 2                      # = b2 - (n-3)/n
    want_q = 3.0 * (n - 1) / (n + 1) - (n - 3) / n
    var_b2 = 24.0 * n * (n - 2) * (n - 3) / ((n + 1) ** 2 * (n + 3) * (n + 5))
    assert abs(q.mean() - want_q) < 5 * np.sqrt(var_b2 / q.size), (q.mean(), want_q)


@pytest.mark.parametrize("n,pairs", CASES)
def test_bonett_equality_test_rejects_at_its_level(n, pairs):
    lnS2, se2 = _simulate(n, pairs)
    a, b = lnS2[:pairs] - lnS2[pairs:], se2[:pairs] + se2[pairs:]
    rate = np.count_nonzero(a * a > Z * Z * b) / pairs
    assert 0.0035 <= rate <= 0.0075, rate


# ------------------------------------------------------- weights by hand
def _accept_all_film(W, H, film):
    """Welch inputs under which every pair is accepted: equal test means, a huge variance."""
    n = np.full((H, W), 16, np.int32)
    mean = np.repeat(film[None], 1, axis=0).astype(np.float32)
    var = np.full((1, 3, H, W), 1e6, np.float32)
    return n, mean, var


def _accept_all_vars(W, H, m2):
    n = np.full((H, W), 9, np.int32)
    lnS2 = np.zeros((1, 3, H, W), np.float32)
    se2 = np.ones((1, 3, H, W), np.float32)
    return n, m2[None].astype(np.float32), lnS2, se2


def test_weights_1x3_by_hand():
    sd_s, sd_g = 1.5, 0.1
    ds = -0.5 / sd_s ** 2
    film = np.array([[1.0, 2.0, 4.0], [3.0, 0.5, 1.0], [7.0, 7.0, 1.0]], np.float32).reshape(3, 1, 3)
    g = np.array([0.0, sd_g, 2 * sd_g], np.float32).reshape(1, 1, 3)   # neighbours one and two sds apart
    e1 = np.exp(ds * 1 - 0.5)     # |dx| = 1, |dg| = 1 sd
    e2 = np.exp(ds * 4 - 2.0)     # |dx| = 2, |dg| = 2 sd
    want = np.empty((3, 1, 3))
    f = film[:, 0].astype(np.float64)
    want[:, 0, 0] = (f[:, 0] + e1 * f[:, 1] + e2 * f[:, 2]) / (1 + e1 + e2)
    want[:, 0, 1] = (f[:, 1] + e1 * (f[:, 0] + f[:, 2])) / (1 + 2 * e1)
    want[:, 0, 2] = (f[:, 2] + e1 * f[:, 1] + e2 * f[:, 0]) / (1 + e1 + e2)
    n, mean, var = _accept_all_film(3, 1, film)
    out = oracle.denoise_film(n, mean, var, film, 2, ds, 0.005, feats=[(g, sd_g)])
    np.testing.assert_allclose(out["filmD"], want, rtol=1e-6)  # features are stored in FP32
    # the variance filter weights its neighbours' s^2 = m2/(n-1) the same way
    n9, m2, lnS2, se2 = _accept_all_vars(3, 1, film * 8.0)
    vD, vCD = oracle.denoise_vars(n9, m2, lnS2, se2, m2 * 0.5, 2, ds, Z, feats=[(g, sd_g)])
    np.testing.assert_allclose(vD[0], want, rtol=1e-6)
    np.testing.assert_allclose(vCD[0], 0.5 * want, rtol=1e-6)


def test_weights_3x3_two_features_by_hand():
    sd_s, sd_a, sd_d = 2.0, 0.02, 0.5
    ds = -0.5 / sd_s ** 2
    rng = np.random.default_rng(3)
    film = rng.uniform(0.1, 5.0, (3, 3, 3)).astype(np.float32)
    alb = np.full((3, 3, 3), 0.4, np.float32)
    dep = np.full((1, 3, 3), 10.0, np.float32)
    alb[0, 0, 0] += sd_a
    alb[1, 0, 0] += sd_a           # (0,0): two albedo channels one sd away: exp(-1)
    dep[0, 1, 2] += sd_d           # (1,2): depth one sd away: exp(-1/2)
    alb[2, 2, 2] -= sd_a
    dep[0, 2, 2] += 2 * sd_d       # (2,2): albedo one sd and depth two sds away: exp(-1/2 - 2)
    # weights of the centre (1,1): exp(ds (dx^2 + dy^2)) times the feature factor
    expo = {(0, 0): 2 * ds - 1.0, (0, 1): ds, (0, 2): 2 * ds,
            (1, 0): ds, (1, 1): 0.0, (1, 2): ds - 0.5,
            (2, 0): 2 * ds, (2, 1): ds, (2, 2): 2 * ds - 2.5}
    w = np.array([[np.exp(expo[(y, x)]) for x in range(3)] for y in range(3)])
    want = (film.astype(np.float64) * w).sum(axis=(1, 2)) / w.sum()
    n, mean, var = _accept_all_film(3, 3, film)
    feats = [(alb, sd_a), (dep, sd_d)]
    out = oracle.denoise_film(n, mean, var, film, 1, ds, 0.005, feats=feats)
    np.testing.assert_allclose(out["filmD"][:, 1, 1], want, rtol=1e-6)  # features are stored in FP32
    n9, m2, lnS2, se2 = _accept_all_vars(3, 3, film * 8.0)
    vD, _ = oracle.denoise_vars(n9, m2, lnS2, se2, m2, 1, ds, Z, feats=feats)
    np.testing.assert_allclose(vD[0][:, 1, 1], want, rtol=1e-6)


# ----------------------------------------------------------------- mode 3
def _two_sets():
    """Pixels 0 and 1 of a 2x1 film: set 0 accepts the pair (close means, wide
    variance), set 1 rejects it (far means, narrow variance)."""
    n = np.full((1, 2), 16, np.int32)
    mean = np.zeros((2, 3, 1, 2), np.float32)
    var = np.zeros((2, 3, 1, 2), np.float32)
    mean[0, :, 0] = [1.0, 1.01]
    var[0] = 1.0
    mean[1, :, 0] = [1.0, 50.0]
    var[1] = 1e-4
    film = np.zeros((3, 1, 2), np.float32)
    film[:, 0] = [2.0, 6.0]
    return n, mean, var, film


@pytest.mark.parametrize("swap", [False, True])
def test_mode3_tests_only_the_smallest_jb_set_of_the_centre(swap):
    n, mean, var, film = _two_sets()
    jb = np.array([[[0.1, 5.0]], [[5.0, 0.1]]], np.float32)  # pixel 0 picks set 0, pixel 1 set 1
    if swap:
        jb = jb[::-1].copy()                                   # pixel 0 picks set 1, pixel 1 set 0
    out = oracle.denoise_film(n, mean, var, film, 1, 0.0, 0.005, mt_mode=3, jb=jb, want_decisions=True)
    avg = 4.0
    if not swap:
        np.testing.assert_allclose(out["filmD"][:, 0], [[avg, 6.0]] * 3)
    else:
        np.testing.assert_allclose(out["filmD"][:, 0], [[2.0, avg]] * 3)
    dec = out["decisions"]  # [H, W, win, win, 3]; pixel 0's neighbour (dx = +1), pixel 1's (dx = -1)
    assert np.all(dec[0, 0, 1, 2] == (0 if swap else 1)) and np.all(dec[0, 1, 1, 0] == (1 if swap else 0))
    # mode 2 needs every set to accept: nothing is averaged
    out2 = oracle.denoise_film(n, mean, var, film, 1, 0.0, 0.005, mt_mode=2, jb=jb)
    np.testing.assert_allclose(out2["filmD"][:, 0], [[2.0, 6.0]] * 3)
