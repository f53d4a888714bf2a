"""Gaussian angle proposals with per-temperature widths for the XY model.

The reference draws a uniform window theta + delta (2u - 1)
(models/xy/update.py:24-31).  ``XYModel(proposal="gaussian")`` instead draws
theta + sigma_t z with z standard normal, where sigma_t belongs to the ladder
temperature t of the walker: the plain path takes z by Box-Muller from two
uniforms, the batched engines by inverse-CDF sampling from one 64-bit draw
(a table of the normal quantile).  The proposal is symmetric, so the
reference's Metropolis acceptance is unchanged.

Widths are adapted toward a target acceptance (0.5 by default) during
thermalisation only -- ``adapt`` -- and are then frozen, so the measured
chain is a fixed-kernel Metropolis chain:

    log sigma_t += rate * (acceptance_t - target),  sigma_t in [1e-4, 4 pi].

Walkers find their temperature through their beta (the ladder beta nearest
to it), so parallel tempering keeps exchanging betas only.
"""

import weakref

import numpy as np

from mcgpu import device_zeros, to_device

SIGMA_MIN, SIGMA_MAX = 1e-4, 4 * np.pi


class GaussianTables:
    """Device tables of one ladder: betas, sigmas and accepted-move counters."""

    def __init__(self, temperatures, L, sigma0, dtype):
        temperatures = np.asarray(temperatures, dtype=np.float64)
        self.dtype = np.dtype(dtype)
        self.L = int(L)
        self.T = len(temperatures)
        self.table_betas = to_device((1.0 / temperatures).astype(self.dtype))
        sig = np.broadcast_to(np.asarray(sigma0, dtype=np.float64), (self.T,))
        if not np.all((sig > 0) & np.isfinite(sig)):
            raise ValueError("sigma must be positive and finite")
        self.sigmas_host = sig.copy()
        self.sigmas = to_device(self.sigmas_host.astype(self.dtype))
        self.accepted = device_zeros(self.T, np.uint64)
        self.attempts = np.zeros(self.T, dtype=np.int64)

    def count_colour(self):
        """One colour pass: L*L/4 attempts at every temperature."""
        self.attempts += self.L * self.L // 4

    def count_sweeps(self, n):
        """n whole sweeps: L*L attempts per sweep at every temperature."""
        self.attempts += n * self.L * self.L

    def acceptance(self):
        acc = self.accepted.copy_to_host().astype(np.float64)
        return np.divide(acc, self.attempts, out=np.full(self.T, np.nan), where=self.attempts > 0)

    def reset(self):
        self.accepted.fill_zero()
        self.attempts[:] = 0

    def set_sigmas(self, sig):
        sig = np.clip(np.asarray(sig, dtype=np.float64), SIGMA_MIN, SIGMA_MAX)
        self.sigmas_host = sig.copy()
        self.sigmas.copy_to_device(sig.astype(self.dtype))

    def adapt_step(self, target=0.5, rate=1.0):
        """Move every sigma toward the target acceptance; returns the rates used."""
        rates = self.acceptance()
        ok = np.isfinite(rates)
        sig = self.sigmas_host.copy()
        sig[ok] *= np.exp(rate * (rates[ok] - target))
        self.set_sigmas(sig)
        self.reset()
        return rates


class ProposalRegistry:
    """Tables per Runtime, keyed weakly by the runtime's betas array."""

    def __init__(self):
        self._tables = weakref.WeakKeyDictionary()

    def attach(self, betas, tables):
        self._tables[betas] = tables

    def get(self, betas):
        tables = self._tables.get(betas)
        if tables is None:
            raise ValueError("Gaussian XY proposals need the model's temperature tables: create the Runtime "
                             "with this model (Runtime attaches them)")
        return tables


def adapt(sweeper, tables, sweeps, interval=50, target=0.5, rate=1.0):
    """Thermalise with `sweeper.sweep()` while adapting sigmas every `interval` sweeps.

    `sweeper` is a Runtime or a ParallelTempering.  Returns the history of
    (sigmas, acceptance) per adaptation step.
    """
    if interval < 1:
        raise ValueError("interval must be positive")
    history = []
    tables.reset()
    done = 0
    while done < sweeps:
        n = min(interval, sweeps - done)
        for _ in range(n):
            sweeper.sweep()
        done += n
        rates = tables.adapt_step(target, rate)
        history.append((tables.sigmas_host.copy(), rates))
    tables.reset()
    return history
