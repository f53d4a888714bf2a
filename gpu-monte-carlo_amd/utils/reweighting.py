"""Float64 histogram reweighting of energy samples.

The reference records per-temperature samples of the energy and the other
observables (core/simulation.py:36-72; with parallel tempering the samples
are filed by temperature, core/parallel_tempering.py:271-273) and estimates
every observable at the simulated temperatures only (utils/Analysis/*,
models/*/analysis.py).  Reweighting moves those estimates to other inverse
temperatures of the same Hamiltonian:

* single histogram (Ferrenberg & Swendsen, PRL 61, 2635 (1988)): samples
  (E_i, O_i) of one chain at beta_0 give
      <O>_beta = sum_i O_i w_i / sum_i w_i,   w_i = exp(-(beta - beta_0) E_i);
* multiple histograms (Ferrenberg & Swendsen, PRL 63, 1195 (1989)): the
  chains at every ladder temperature beta_k (n_k samples each) are combined
  with the weight of a sample of energy E
      W_beta(E) = exp(-beta E) / sum_k n_k exp(-beta_k E + f_k),
  where the free energies f_k = log Z(beta_k) solve the fixed point
      f_j = log sum_{all samples} W_{beta_j}(E)    (up to one common constant).

Everything is float64 and every sum of exponentials is a log-sum-exp
(shifted by its largest term), so no weight overflows however far the
energies spread.  Errors come from a delete-one-block jackknife over time
blocks of the samples (utils/analysis.py), with the free energies solved
again for every deleted block.
"""

import numpy as np


def _logsumexp(a, axis=None, b=None):
    a = np.asarray(a, dtype=np.float64)
    top = np.max(a, axis=axis, keepdims=True)
    top = np.where(np.isfinite(top), top, 0.0)
    e = np.exp(a - top)
    if b is not None:
        e = e * b
    s = np.sum(e, axis=axis, keepdims=True)
    out = np.log(s) + top
    return np.squeeze(out, axis=axis) if axis is not None else float(out.squeeze())


def single_histogram(energies, observables, beta0, betas, weights=None):
    """<O>_beta for every beta in `betas` from samples of one chain at beta0.

    energies: (n,) sample energies; observables: {name: (n,) values} (the
    energy itself is always included as "energy"); weights: optional (n,)
    sample multiplicities.  Returns {name: (len(betas),) float64}."""
    E = np.asarray(energies, dtype=np.float64).ravel()
    betas = np.atleast_1d(np.asarray(betas, dtype=np.float64))
    w = np.ones_like(E) if weights is None else np.asarray(weights, dtype=np.float64).ravel()
    logw = -(betas[:, None] - float(beta0)) * E[None, :]               # (B, n)
    logw = logw + np.log(np.where(w > 0, w, 1.0))[None, :] + np.where(w > 0, 0.0, -np.inf)[None, :]
    norm = _logsumexp(logw, axis=1)
    p = np.exp(logw - norm[:, None])                                   # normalised weights per beta
    obs = {"energy": E, **{k: np.asarray(v, dtype=np.float64).ravel() for k, v in observables.items()}}
    return {k: p @ v for k, v in obs.items()}


class MultiHistogram:
    """Multiple-histogram estimator over the chains of a temperature ladder.

    energies: (n, R) samples, column k taken at inverse temperature betas[k]
    (a run's samples in temperature order); observables: {name: (n, R)}.
    """

    def __init__(self, energies, betas, observables=None, tol=1e-13, max_iter=100000):
        self.E = np.asarray(energies, dtype=np.float64)
        if self.E.ndim != 2:
            raise ValueError("energies must be (samples, temperatures)")
        self.betas = np.asarray(betas, dtype=np.float64)
        if self.betas.shape != (self.E.shape[1],):
            raise ValueError("one beta per energy column")
        self.obs = {k: np.asarray(v, dtype=np.float64) for k, v in (observables or {}).items()}
        for k, v in self.obs.items():
            if v.shape != self.E.shape:
                raise ValueError(f"observable {k} must have the energies' shape")
        self.tol, self.max_iter = tol, max_iter
        self.f = self._solve(self.E)

    def _solve(self, E, start=None):
        """f_k = log Z(beta_k) - log Z(beta_0) of the energies shifted by their mean, by
        iterating the self-consistent equations from `start` (zeros by default)."""
        flat = E.ravel()
        n = np.full(len(self.betas), E.shape[0], dtype=np.float64)
        f = np.zeros(len(self.betas)) if start is None else np.array(start, dtype=np.float64)
        # shift the energies so exp(-beta E) stays in range: only differences matter
        shift = flat.mean()
        x = flat - shift
        a = -self.betas[:, None] * x[None, :]                           # (R, N): -beta_k E
        for _ in range(self.max_iter):
            # log denominator per sample: log sum_k n_k exp(-beta_k E + ... - f_k)
            den = _logsumexp(a - f[:, None], axis=0, b=n[:, None])
            new = _logsumexp(a - den[None, :], axis=1)
            new -= new[0]
            if np.max(np.abs(new - f)) < self.tol:
                f = new
                break
            f = new
        return f

    def _log_weights(self, E, f, beta):
        x = E.ravel() - E.ravel().mean()
        n = np.full(len(self.betas), E.shape[0], dtype=np.float64)
        den = _logsumexp(-self.betas[:, None] * x[None, :] - f[:, None], axis=0, b=n[:, None])
        return -beta * x - den

    def expectations(self, betas, E=None, obs=None, f=None):
        """{name: <O>_beta} for every beta, plus "energy", "energy2" and "log_z"
        (log Z(beta) - log Z(betas[0]))."""
        E = self.E if E is None else E
        obs = self.obs if obs is None else obs
        f = self.f if f is None else f
        betas = np.atleast_1d(np.asarray(betas, dtype=np.float64))
        values = {"energy": E.ravel(), "energy2": E.ravel()**2, **{k: v.ravel() for k, v in obs.items()}}
        out = {k: np.empty(len(betas)) for k in values}
        out["log_z"] = np.empty(len(betas))
        # the estimator sums weights of the mean-shifted energies; exp(-beta shift) undoes the shift
        shift = E.ravel().mean()
        base = _logsumexp(self._log_weights(E, f, self.betas[0])) - self.betas[0] * shift
        for i, b in enumerate(betas):
            lw = self._log_weights(E, f, b)
            norm = _logsumexp(lw)
            p = np.exp(lw - norm)
            for k, v in values.items():
                out[k][i] = p @ v
            out["log_z"][i] = norm - b * shift - base
        return out

    def log_partition_ratio(self, beta):
        """log Z(beta) - log Z(betas[0])."""
        return float(self.expectations([beta])["log_z"][0])

    def jackknife(self, betas, block_size):
        """Delete-one-block jackknife of expectations(betas): (values, errors) dicts,
        the free energies solved again without each block."""
        n = self.E.shape[0] // block_size
        if n < 2:
            raise ValueError("need at least two complete time blocks")
        full = self.expectations(betas)
        reps = []
        for b in range(n):
            keep = np.ones(self.E.shape[0], dtype=bool)
            keep[b * block_size:(b + 1) * block_size] = False
            keep[n * block_size:] = False
            E = self.E[keep]
            obs = {k: v[keep] for k, v in self.obs.items()}
            reps.append(self.expectations(betas, E, obs, self._solve(E, self.f)))
        errors = {k: np.sqrt((n - 1) * np.var([r[k] for r in reps], axis=0)) for k in full}
        return full, errors


def reweight_run(run, betas, observables=None, block_size=None):
    """Multiple-histogram reweighting of a run dictionary (temperature-ordered
    samples, reference layout) to inverse temperatures `betas`.  Returns
    {name: values} and, with block_size, {name + "_err": jackknife errors};
    "specific_heat" per site is beta^2 (<E^2> - <E>^2) / N with N = L^2."""
    s, md = run["samples"], run["metadata"]
    T = np.asarray(md["temperatures"], dtype=np.float64)
    names = list(observables or [])
    mh = MultiHistogram(s["energy"], 1.0 / T, {k: s[k] for k in names})
    betas = np.atleast_1d(np.asarray(betas, dtype=np.float64))
    N = float(md["L"]) ** 2
    if block_size is None:
        vals = mh.expectations(betas)
        errs = None
    else:
        vals, errs = mh.jackknife(betas, block_size)
    out = {k: v for k, v in vals.items()}
    out["specific_heat"] = betas**2 * (vals["energy2"] - vals["energy"]**2) / N
    if errs is not None:
        for k, e in errs.items():
            out[k + "_err"] = e
    return out
