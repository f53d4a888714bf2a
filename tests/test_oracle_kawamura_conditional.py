"""Pin of the oracle's Kawamura sampler by its exact single-site conditional.

The reference's update (models/kawamura/update.py:21-65) runs three
Metropolis sub-moves per visit -- a window move of theta, a flip of tau and
an independent uniform draw of the unit vector d -- each accepted against the
energy left by the previous sub-move.  Every proposal is symmetric with
respect to the measure dtheta x counting(tau) x dOmega(d), so a visit leaves
the single-site conditional

    pi(theta, tau, d)  ~  exp(-beta E_loc),
    E_loc = -sum_n J (1 + tau tau_n) (d . d_n) cos(theta - theta_n)

invariant.  With only colour 0 updated, the neighbours (colours 1 and 2)
never change and every colour-0 site is a chain with that stationary law.

The law has closed forms: for fixed (theta, tau), E_loc = -d . h with
h = sum_n 2 J [tau_n = tau] cos(theta - theta_n) d_n, so d is von
Mises-Fisher with kappa = beta |h|; integrating d over the sphere leaves the
weight sinh(kappa) / kappa on (theta, tau), and <d | theta, tau> =
h_hat (coth kappa - 1 / kappa).  One periodic quadrature over theta then gives
<tau>, <cos 2 theta>, <sin 2 theta>, <d . h> and <d cos theta>, <d sin theta>.
E_loc is unchanged by (theta, d) -> (theta + pi, -d), so the first moments
<cos theta>, <sin theta> and <d> vanish.  At large beta a chain stays in one
of the two mirror images (a window move cannot carry theta across while d
flips), so those odd moments are checked only at the two smaller betas, where
the chains cross between the images; the other moments, which the mirror
leaves unchanged, are checked at every beta.  (A sampler that leaves half of
the sphere unproposed is seen only by the odd moments: its law, added to its
mirror image, is the right one.)

The same check is run on a numpy restatement with planted mistakes (a stale
energy in the tau or d sub-move, a dropped (1 + tau tau_n) weight, a
non-uniform sphere proposal, a halved acceptance exponent): each must fail
it, so the check has the power to see them.
"""

import numpy as np
import pytest

import oracle

J = 1.0
L = 8
BETAS = np.array([0.5, 1.0, 1.6, 2.4])
ODD_BETA_MAX = 1.0  # odd moments (zero by the mirror symmetry) are checked up to this beta
N_EVEN, N_ODD = 10, 5
SITES = [(i, j) for i in range(0, L, 2) for j in range(0, L, 2)]  # colour 0


def _neighbour_sites(i, j):
    return [((i - 1) % L, j), ((i + 1) % L, j), (i, (j - 1) % L), (i, (j + 1) % L)]


def _frozen_state(seed):
    g = np.random.default_rng(seed)
    R = len(BETAS)
    st = np.zeros((R, L, L, 5))
    st[..., 0] = g.uniform(-np.pi, np.pi, (R, L, L))
    st[..., 1] = g.choice([-1.0, 1.0], (R, L, L))
    z = g.uniform(-1, 1, (R, L, L))
    phi = g.uniform(0, 2 * np.pi, (R, L, L))
    rad = np.sqrt(1 - z * z)
    st[..., 2], st[..., 3], st[..., 4] = rad * np.cos(phi), rad * np.sin(phi), z
    return st


def _neighbours(st):
    """(R, sites, 4, 5) neighbour records of every colour-0 site."""
    return np.array([[[st[r, a, b] for a, b in _neighbour_sites(i, j)] for (i, j) in SITES]
                     for r in range(len(BETAS))])


def _field(theta, tau, nb):
    """h(theta, tau) for neighbour records nb (4, 5); theta an array."""
    on = (nb[:, 1] == tau).astype(float)                       # (1 + tau tau_n) / 2
    w = 2 * J * on[None, :] * np.cos(theta[:, None] - nb[None, :, 0])   # (n, 4)
    return w @ nb[:, 2:5]                                       # (n, 3)


def exact_moments(nb, beta, n=4096):
    """<tau>, <cos 2 theta>, <sin 2 theta>, <d.h>, <d cos theta> (3), <d sin theta> (3)
    of the conditional."""
    theta = -np.pi + 2 * np.pi * np.arange(n) / n
    logw, acc = [], []
    for tau in (-1.0, 1.0):
        h = _field(theta, tau, nb)
        hn = np.linalg.norm(h, axis=1)
        k = beta * hn
        small = k < 1e-8
        ks = np.where(small, 1.0, k)
        lw = np.where(small, 0.0, ks + np.log1p(-np.exp(-2 * ks)) - np.log(2 * ks))
        lang = np.where(small, k / 3, 1 / np.tanh(ks) - 1 / ks)  # <d . h_hat>
        hhat = h / np.where(hn > 0, hn, 1.0)[:, None]
        dmean = hhat * lang[:, None]
        obs = np.stack([np.full(n, tau), np.cos(2 * theta), np.sin(2 * theta), hn * lang,
                        *(dmean * np.cos(theta)[:, None]).T, *(dmean * np.sin(theta)[:, None]).T], axis=1)
        logw.append(lw)
        acc.append(obs)
    lw = np.concatenate(logw)
    w = np.exp(lw - lw.max())
    obs = np.concatenate(acc)
    even = (w[:, None] * obs).sum(axis=0) / w.sum()
    return np.concatenate([even, np.zeros(N_ODD)])  # odd moments: zero by the mirror symmetry


def sample_observables(samples, nb):
    """Per step, the observables of exact_moments for one chain."""
    theta, tau, d = samples[:, 0], samples[:, 1], samples[:, 2:5]
    dh = np.empty(len(samples))
    for t in (-1.0, 1.0):
        sel = tau == t
        if sel.any():
            dh[sel] = np.einsum("ij,ij->i", _field(theta[sel], t, nb), d[sel])
    c, s = np.cos(theta)[:, None], np.sin(theta)[:, None]
    return np.concatenate([np.stack([tau, np.cos(2 * theta), np.sin(2 * theta), dh], axis=1), d * c, d * s,
                           c, s, d], axis=1)


def _binned(x, nbins=20):
    n = len(x) // nbins * nbins
    b = x[:n].reshape(nbins, -1, *x.shape[1:]).mean(axis=1)
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(nbins)


def z_scores(trace, nbs):
    """z of every (chain, observable) against the exact moments; trace (steps, R, sites, 5).
    Odd moments of chains above ODD_BETA_MAX are NaN (not checked)."""
    out = []
    for r, beta in enumerate(BETAS):
        for s in range(len(SITES)):
            obs = sample_observables(trace[:, r, s], nbs[r, s])
            mean, err = _binned(obs)
            z = (mean - exact_moments(nbs[r, s], beta)) / np.maximum(err, 1e-4)
            if beta > ODD_BETA_MAX:
                z[N_EVEN:] = np.nan
            out.append(z)
    return np.asarray(out)


def _summary(z):
    """Largest |z|, and the largest mean z^2 over chains of one observable (a
    bias shared by many chains shows there before any single z is large)."""
    return float(np.nanmax(np.abs(z))), float(np.max(np.nanmean(z**2, axis=0)))


def test_exact_moments_limits():
    """beta -> 0: the conditional is uniform (all moments zero).  A single
    aligned neighbour and tau = tau_n forced: <d . h> is 2J <cos> times the
    Langevin mean; checked against brute-force 3-D quadrature."""
    nb = np.zeros((4, 5))
    nb[:, 1] = 1.0
    nb[:, 4] = 1.0
    nb[:, 0] = [0.3, -1.1, 2.0, 0.0]
    m0 = exact_moments(nb, 1e-9)
    np.testing.assert_allclose(m0, 0.0, atol=1e-6)
    # brute force over theta, tau and the sphere (z, phi grid)
    beta = 1.3
    th = -np.pi + 2 * np.pi * np.arange(256) / 256
    z = np.polynomial.legendre.leggauss(64)
    phi = 2 * np.pi * np.arange(64) / 64
    Z = tot = ttau = 0.0
    for tau in (-1.0, 1.0):
        for zz, wz in zip(*z):
            rad = np.sqrt(1 - zz * zz)
            d = np.stack([rad * np.cos(phi), rad * np.sin(phi), np.full(64, zz)], 1)  # (64, 3)
            h = _field(th, tau, nb)                                                  # (256, 3)
            e = -(h @ d.T)                                                           # (256, 64)
            w = wz * np.exp(-beta * e)
            Z += w.sum()
            tot += (w * -e).sum()
            ttau += tau * w.sum()
    m = exact_moments(nb, beta)
    np.testing.assert_allclose(m[3], tot / Z, rtol=1e-9)
    np.testing.assert_allclose(m[0], ttau / Z, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_kawamura_samples_the_exact_conditional(dtype):
    st = _frozen_state(31).astype(dtype)
    nbs = _neighbours(st.astype(float))
    rng = oracle.rng_init(st.shape[0] * L * L, 77)
    betas = BETAS.astype(dtype)
    frozen = st.copy()
    for _ in range(2000):
        oracle.kawamura_color(st, rng, betas, 0, J, 1.5)
    steps = 40000
    trace = np.empty((steps, len(BETAS), len(SITES), 5))
    ii = np.array([i for i, _ in SITES])
    jj = np.array([j for _, j in SITES])
    for n in range(steps):
        oracle.kawamura_color(st, rng, betas, 0, J, 1.5)
        trace[n] = st[:, ii, jj, :]
    other = np.ones((L, L), bool)
    other[ii, jj] = False
    np.testing.assert_array_equal(st[:, other], frozen[:, other])  # neighbours stay frozen
    worst, per_obs = _summary(z_scores(trace, nbs))
    assert worst < 5.5 and per_obs < 2.5, (worst, per_obs)


# ------------------------------------------------------------------ power
MISTAKES = (None, "stale_tau", "stale_d", "dropped_weight", "half_circle", "half_beta")


def _numpy_chains(nbs, steps, seed, delta=1.5):
    """Restatement of the three sub-moves (float64, numpy's generator) for
    every chain, once per entry of MISTAKES: variant 0 is the reference rule,
    the others each carry one planted mistake."""
    g = np.random.default_rng(seed)
    V = len(MISTAKES)
    R, S = nbs.shape[:2]
    shape = (V, R, S)
    is_ = {m: np.array([x == m for x in MISTAKES])[:, None, None] for m in MISTAKES[1:]}
    beta = BETAS[None, :, None]
    th = np.zeros(shape)
    tau = np.ones(shape)
    d = np.zeros(shape + (3,))
    d[..., 2] = 1.0
    nb = nbs[None]
    ntheta, ntau = nb[..., 0], nb[..., 1]
    nx, ny, nz = nb[..., 2], nb[..., 3], nb[..., 4]
    plain = is_["dropped_weight"][..., None]

    def energy(t_, ta, dd):
        w = np.where(plain, 1.0, 1 + ta[..., None] * ntau)
        dot = dd[..., 0, None] * nx + dd[..., 1, None] * ny + dd[..., 2, None] * nz
        return -(J * w * dot * np.cos(t_[..., None] - ntheta)).sum(-1)

    def accept(dE, factor=1.0):
        return (dE <= 0) | (g.random(shape) < np.exp(-factor * beta * dE))

    out = np.empty((steps,) + shape + (5,))
    for n in range(steps):
        e = e_before = energy(th, tau, d)
        p = th + delta * (2 * g.random(shape) - 1)
        p = np.where(p >= np.pi, p - 2 * np.pi, np.where(p < -np.pi, p + 2 * np.pi, p))
        pe = energy(p, tau, d)
        a = accept(pe - e)
        th = np.where(a, p, th)
        e = np.where(a, pe, e)
        pe = energy(th, -tau, d)
        a = accept(pe - np.where(is_["stale_tau"], e_before, e), np.where(is_["half_beta"], 0.5, 1.0))
        e_tau = e
        tau = np.where(a, -tau, tau)
        e = np.where(a, pe, e)
        z = 2 * g.random(shape) - 1
        phi = np.where(is_["half_circle"], np.pi, 2 * np.pi) * g.random(shape)
        rad = np.sqrt(1 - z * z)
        x = np.stack([rad * np.cos(phi), rad * np.sin(phi), z], -1)
        pe = energy(th, tau, x)
        a = accept(pe - np.where(is_["stale_d"], e_tau, e))
        d = np.where(a[..., None], x, d)
        out[n, ..., 0], out[n, ..., 1], out[n, ..., 2:5] = th, tau, d
    return out[steps // 20:]


@pytest.fixture(scope="module")
def planted():
    nbs = _neighbours(_frozen_state(31))
    return nbs, _numpy_chains(nbs, 30000, 5)


@pytest.mark.parametrize("variant", range(len(MISTAKES)), ids=[str(m) for m in MISTAKES])
def test_conditional_check_detects_planted_mistakes(planted, variant):
    nbs, chains = planted
    worst, per_obs = _summary(z_scores(chains[:, variant], nbs))
    if MISTAKES[variant] is None:
        assert worst < 5.5 and per_obs < 2.5, (worst, per_obs)
    else:
        assert worst > 8.0 or per_obs > 4.0, (MISTAKES[variant], worst, per_obs)
