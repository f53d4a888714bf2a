"""Parallel-tempering diagnostics of saved runs: swap acceptance per ladder
edge, completed walker round trips and walker temperature histories.

Follows utils/Analysis/parallel_tempering.py:7-120 of the reference: the
summary fields and their conventions (rates weighted by attempts, ``None``
for an edge never attempted or a walker with no finished trip, trip means
over finished trips only), the history validation, and the heatmap geometry
(the last assignment kept where sweeps repeat, cell edges at the midpoints
of the recorded sweeps, half a sweep either side of a single record).

The input is what ``ParallelTempering.diagnostics()`` and
``BatchedRun.diagnostics()`` store under ``metadata["parallel_tempering"]``,
as returned by ``utils.data_io.load_runs``.  Runs without that entry are
skipped.  Plotting needs matplotlib; the numbers do not.
"""

import numpy as np

_SCOPE_NOTE = ("Trip means exclude unfinished trips; "
               "zero trips does not imply zero one-way crossings.")


def _ratio(num, den):
    return float(num / den) if den else None


def _checked_history(L, history, n_temperatures):
    ids = np.asarray(history["walker_ids"], dtype=np.int64)
    sweeps = np.asarray(history["sweeps"], dtype=np.int64)
    positions = np.asarray(history["temperature_indices"])
    ok = (ids.ndim == 1 and len(ids) > 0 and len(np.unique(ids)) == len(ids)
          and ids.min() >= 0 and ids.max() < n_temperatures
          and sweeps.ndim == 1 and len(sweeps) > 0 and not np.any(np.diff(sweeps) < 0)
          and positions.shape == (len(sweeps), len(ids)) and np.issubdtype(positions.dtype, np.integer)
          and positions.min() >= 0 and positions.max() < n_temperatures)
    if not ok:
        raise ValueError(f"L={L}: invalid walker history")
    return ids, sweeps, positions.astype(np.int64)


def _edge_summary(T, pt):
    attempts = np.asarray(pt["attempts"])
    accepted = np.asarray(pt["accepted"])
    trips = np.asarray(pt["walker_round_trips"])
    trip_sweeps = np.asarray(pt["walker_round_trip_total_sweeps"])
    return {
        "temperatures": [float(t) for t in T],
        "attempts": attempts.tolist(),
        "accepted": accepted.tolist(),
        "acceptance_rates": [_ratio(a, n) for a, n in zip(accepted, attempts)],
        "aggregate_acceptance": _ratio(accepted.sum(), attempts.sum()),
        "walker_round_trips": trips.tolist(),
        "walker_mean_round_trip_sweeps": [_ratio(t, n) for t, n in zip(trip_sweeps, trips)],
        "mean_round_trip_sweeps": _ratio(trip_sweeps.sum(), trips.sum()),
        "diagnostics_scope": pt["diagnostics_scope"],
    }


class PTAnalysis:
    """Diagnostics of every run in ``{L: run}`` (or ``{(L, i): run}``) that carries PT metadata."""

    def __init__(self, runs):
        self.results = {}
        self._histories = {}
        self._thermalization = {}
        # sizes in order; runs of one size under several models carry (L, i) keys
        for L in sorted(runs, key=lambda k: k if isinstance(k, tuple) else (k,)):
            metadata = runs[L].get("metadata", {})
            pt = metadata.get("parallel_tempering")
            if pt is None:
                continue
            T = np.asarray(metadata["temperatures"], dtype=float)
            history = pt.get("walker_history")
            if history is not None:
                self._histories[L] = _checked_history(L, history, len(T))
            self._thermalization[L] = metadata.get("thermalization_sweeps")
            result = _edge_summary(T, pt)
            result["history_available"] = history is not None
            self.results[L] = result

    def summary(self):
        return self.results

    def format_summary(self):
        lines = ["Parallel tempering diagnostics (including thermalization)", _SCOPE_NOTE]
        for L, res in self.results.items():
            lines.append(f"L={L}: {sum(res['accepted'])}/{sum(res['attempts'])} accepted, "
                         f"{sum(res['walker_round_trips'])} trips, "
                         f"mean {res['mean_round_trip_sweeps']} sweeps/trip")
            lines.append("  walker paths recorded" if res["history_available"] else
                         "  walker paths were not recorded; final mappings cannot reconstruct trajectories")
        return "\n".join(lines)

    def walker_heatmap(self, L):
        """Heatmap data of one size's recorded walker paths, or None.

        Returns ``(walker_ids, time_edges, temperatures)``: walkers sorted by
        id; ``temperatures[w, k]`` is the temperature of walker w in the k-th
        distinct recorded sweep (the last record of a repeated sweep wins);
        ``time_edges`` has one more entry than there are distinct sweeps.
        """
        if L not in self._histories:
            return None
        T = np.asarray(self.results[L]["temperatures"])
        ids, sweeps, positions = self._histories[L]
        order = np.argsort(ids)
        ids, positions = ids[order], positions[:, order]
        last = np.append(sweeps[1:] != sweeps[:-1], True)
        times = sweeps[last].astype(float)
        if len(times) == 1:
            edges = np.array([times[0] - 0.5, times[0] + 0.5])
        else:
            edges = np.concatenate([[times[0]], 0.5 * (times[1:] + times[:-1]), [times[-1]]])
        return ids, edges, T[positions[last]].T

    def plot(self, show=True):
        """Swap acceptance per edge midpoint and round trips per walker."""
        if not self.results:
            return None
        import matplotlib.pyplot as plt
        fig, (acc_ax, trip_ax) = plt.subplots(1, 2, figsize=(12, 5))
        for L, res in self.results.items():
            T = np.asarray(res["temperatures"])
            rates = [np.nan if r is None else r for r in res["acceptance_rates"]]
            acc_ax.plot(0.5 * (T[1:] + T[:-1]), rates, ".-", label=f"L={L}")
            trip_ax.plot(res["walker_round_trips"], ".-", label=f"L={L}")
        acc_ax.set(xlabel="Temperature midpoint", ylabel="Swap acceptance", ylim=(0, 1.05))
        trip_ax.set(xlabel="Walker", ylabel="Completed round trips")
        for ax in (acc_ax, trip_ax):
            ax.grid(alpha=0.3)
            ax.legend()
        fig.tight_layout()
        if show:
            plt.show()
        return fig

    def plot_walkers(self, L=None, show=True):
        """Temperature history of every recorded walker as a heatmap, one
        panel per size.  Sampled paths can miss short excursions; the trip
        counters stay authoritative.  Sizes without a history get no panel."""
        sizes = sorted(self._histories) if L is None else ([L] if L in self._histories else [])
        if not sizes:
            return None
        import matplotlib.pyplot as plt
        fig, axes = plt.subplots(len(sizes), 1, figsize=(11, 3.5 * len(sizes)), squeeze=False)
        for size, ax in zip(sizes, axes[:, 0]):
            T = np.asarray(self.results[size]["temperatures"])
            ids, edges, temps = self.walker_heatmap(size)
            mesh = ax.pcolormesh(edges, np.arange(len(ids) + 1) - 0.5, temps, cmap="coolwarm",
                                 vmin=T[0], vmax=T[-1], shading="flat", rasterized=True)
            ticks = np.unique(np.linspace(0, len(ids) - 1, min(7, len(ids))).astype(int))
            ax.set_yticks(ticks, labels=[str(i) for i in ids[ticks]])
            who = "all walkers" if len(ids) == len(T) else f"{len(ids)} recorded walkers"
            ax.set(title=f"L={size}: temperature history of {who} (sampled)",
                   xlabel="Local sweeps since PT construction", ylabel="Walker ID")
            therm = self._thermalization[size]
            sweeps = self._histories[size][1]
            if therm is not None and sweeps[0] <= therm <= sweeps[-1]:
                ax.axvline(therm, color="black", linestyle="--", linewidth=0.8)
            fig.colorbar(mesh, ax=ax, label="Assigned temperature")
        fig.tight_layout()
        if show:
            plt.show()
        return fig
