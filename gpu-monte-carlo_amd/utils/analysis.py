"""Validation estimators for run files: blocked jackknife, Binder cumulants,
helicity modulus, curve crossings and the Weber-Minnhagen BKT fit.

The estimators follow the reference's analysis modules (formulas only; the
reference's plotting is not part of this package):
* blocking and jackknife: utils/Analysis/statistics.py:6-37
* Ising: models/ising/analysis.py:54-75 (energy, |m|, C, chi, Binder with 3)
* XY: models/xy/analysis.py:47-54 (energy, Binder with 2, helicity
  [J<cx> - J^2/T (<qx^2> - <qx>^2)] / L^2)
* Kawamura: models/kawamura/analysis.py:36-74
* He3: models/he3/analysis.py:29-40,103-123
* crossings: utils/Analysis/crossings.py:5-31; WM fit: utils/Analysis/bkt.py:7-48
Each ``*_analysis(run, L, block_size)`` returns ``{name: value, name_err: error}``
arrays over the run's temperatures.
"""

import numpy as np


def block_means(data, block_size):
    data = np.asarray(data)
    n = len(data) // block_size
    return data[:n * block_size].reshape(n, block_size, *data.shape[1:]).mean(axis=1)


def jackknife_from_blocks(blocks, estimator):
    """Delete-one-block jackknife of a (possibly nonlinear) estimator of block means."""
    B = len(next(iter(blocks.values())))
    if B < 2:
        raise ValueError("need at least two complete time blocks")
    totals = {k: v.sum(axis=0) for k, v in blocks.items()}
    estimate = estimator({k: t / B for k, t in totals.items()})
    deleted = np.array([estimator({k: (totals[k] - v[b]) / (B - 1) for k, v in blocks.items()})
                        for b in range(B)])
    return estimate, np.sqrt((B - 1) * np.var(deleted, axis=0, ddof=0))


def blocked_jackknife(moments, estimator, block_size):
    return jackknife_from_blocks({k: block_means(v, block_size) for k, v in moments.items()}, estimator)


def _pack(names, values, errors):
    out = {}
    for name, v, e in zip(names, values, errors):
        out[name] = v
        out[name + "_err"] = e
    return out


def ising_analysis(run, L, block_size=200):
    s, md = run["samples"], run["metadata"]
    T = np.asarray(md["temperatures"], dtype=float)
    N = L * L
    E = np.asarray(s["energy"], float)
    M = np.asarray(s["magnetization"], float)
    moments = {"E": E, "E2": E**2, "M": M, "absM": np.abs(M), "M2": M**2, "M4": M**4}

    def est(m):
        return np.stack((m["E"] / N, m["absM"] / N, (m["E2"] - m["E"]**2) / (N * T**2),
                         (m["M2"] - m["M"]**2) / (N * T), 1 - m["M4"] / (3 * m["M2"]**2)))

    v, e = blocked_jackknife(moments, est, block_size)
    return _pack(("energy", "magnetization", "heat_capacity", "susceptibility", "binder"), v, e)


def xy_analysis(run, L, block_size=200):
    s, md = run["samples"], run["metadata"]
    T = np.asarray(md["temperatures"], dtype=float)
    J = float(md["model_parameters"]["J"])
    mx, my = np.asarray(s["mx"], float), np.asarray(s["my"], float)
    qx = np.asarray(s["twist_sin_x"], float)
    m2 = mx**2 + my**2
    moments = {"energy": np.asarray(s["energy"], float) / L**2, "m2": m2, "m4": m2**2,
               "cx": np.asarray(s["twist_cos_x"], float), "qx": qx, "qx2": qx**2}

    def est(m):
        helicity = (J * m["cx"] - J**2 / T * (m["qx2"] - m["qx"]**2)) / L**2
        return np.stack((m["energy"], 1 - m["m4"] / (2 * m["m2"]**2), helicity))

    v, e = blocked_jackknife(moments, est, block_size)
    return _pack(("energy", "binder", "helicity"), v, e)


def kawamura_analysis(run, L, block_size=100):
    s, md = run["samples"], run["metadata"]
    T = np.asarray(md["temperatures"], dtype=float)
    J = float(md["model_parameters"]["J"])
    tau2 = np.asarray(s["chirality"], float)**2 / L**4
    phase2 = (np.asarray(s["px"], float)**2 + np.asarray(s["py"], float)**2) / L**4
    spin2 = np.asarray(s["spin2"], float) / L**4
    qx = np.asarray(s["twist_sin_x"], float)
    moments = {"energy": np.asarray(s["energy"], float) / L**2, "tau2": tau2, "tau4": tau2**2,
               "phase2": phase2, "phase4": phase2**2, "spin2": spin2, "spin4": spin2**2,
               "cx": np.asarray(s["twist_cos_x"], float), "qx": qx, "qx2": qx**2}

    def est(m):
        helicity = (J * m["cx"] - J**2 / T * (m["qx2"] - m["qx"]**2)) / L**2
        return np.stack((m["energy"], np.sqrt(m["tau2"]), np.sqrt(m["phase2"]), np.sqrt(m["spin2"]),
                         0.5 * (3 - m["tau4"] / m["tau2"]**2), 2 - m["phase4"] / m["phase2"]**2,
                         0.5 * (7 - 5 * m["spin4"] / m["spin2"]**2), helicity))

    v, e = blocked_jackknife(moments, est, block_size)
    return _pack(("energy", "chirality", "phase", "spin", "binder_chirality", "binder_phase",
                  "binder_spin", "helicity"), v, e)


def he3_analysis(run, L, block_size=100):
    s, md = run["samples"], run["metadata"]
    T = np.asarray(md["temperatures"], dtype=float)
    chir = np.asarray(s["chirality"], float) / np.asarray(s["density"], float)
    phase2 = (np.asarray(s["phase_real"], float)**2 + np.asarray(s["phase_imag"], float)**2) / L**4
    ix, iy = np.asarray(s["current_x"], float), np.asarray(s["current_y"], float)
    moments = {"energy": np.asarray(s["energy"], float) / L**2,
               "density": np.asarray(s["density"], float) / L**2,
               "chirality": np.abs(chir), "chi2": chir**2, "chi4": chir**4,
               "phase2": phase2, "phase4": phase2**2, "ix": ix, "ix2": ix**2,
               "kx": np.asarray(s["curvature_x"], float), "iy": iy, "iy2": iy**2,
               "ky": np.asarray(s["curvature_y"], float)}
    names = ("energy", "density", "chirality", "phase", "binder_chirality", "binder_phase",
             "helicity_x", "helicity_y", "helicity")
    if "spin_alignment" in s:
        moments["spin_alignment"] = np.asarray(s["spin_alignment"], float)
        names += ("spin_alignment",)

    def est(m):
        hx = (m["kx"] - (m["ix2"] - m["ix"]**2) / T) / L**2
        hy = (m["ky"] - (m["iy2"] - m["iy"]**2) / T) / L**2
        vals = (m["energy"], m["density"], m["chirality"], np.sqrt(m["phase2"]),
                1 - m["chi4"] / (3 * m["chi2"]**2), 2 - m["phase4"] / m["phase2"]**2,
                hx, hy, 0.5 * (hx + hy))
        if "spin_alignment" in m:
            vals += (m["spin_alignment"],)
        return np.stack(vals)

    v, e = blocked_jackknife(moments, est, block_size)
    return _pack(names, v, e)


def linear_crossings(T, y1, y2=0.0, *, temperature_range=None):
    """Temperatures where two piecewise-linear curves cross ([nan] if ill-defined)."""
    T = np.asarray(T)
    d = np.broadcast_to(np.asarray(y1) - y2, T.shape)
    if not np.all(np.isfinite(d)) or np.any((d[:-1] == 0) & (d[1:] == 0)):
        return [np.nan]
    roots = list(T[d == 0])
    for i in np.where(d[:-1] * d[1:] < 0)[0]:
        roots.append(T[i] - d[i] * (T[i + 1] - T[i]) / (d[i + 1] - d[i]))
    if temperature_range is not None:
        lo, hi = temperature_range
        roots = [t for t in roots if lo <= t <= hi]
    return sorted(float(t) for t in roots)


def fit_weber_minnhagen(T, L, helicity, errors):
    """Fit Upsilon(Tc, L) = 2 Tc / pi (1 + 1 / (2 ln L + C)); arrays (sizes, T)."""
    from scipy.optimize import least_squares
    T, L = np.asarray(T, float), np.asarray(L, float)
    rho, sigma = np.asarray(helicity, float), np.asarray(errors, float)
    log_sizes = 2 * np.log(L / L[0])

    def residual(params):
        Tc, a = params
        measured = np.array([np.interp(Tc, T, row) for row in rho])
        unc = np.array([np.interp(Tc, T, row) for row in sigma])
        return (measured - 2 * Tc / np.pi * (1 + a / (1 + a * log_sizes))) / unc

    fits = [least_squares(residual, [t, 0.1], bounds=([T[0], 0], [T[-1], np.inf]))
            for t in np.linspace(T[0], T[-1], 5)]
    best = min(fits, key=lambda f: float(np.dot(f.fun, f.fun)))
    Tc, a = best.x
    boundary = min(abs(Tc - T[0]), abs(Tc - T[-1])) < 1e-6 * (T[-1] - T[0])
    return {"temperature": None if boundary or a <= 1e-8 else float(Tc),
            "C": float(1 / a - 2 * np.log(L[0])) if a > 1e-8 else None,
            "chi2": float(np.dot(best.fun, best.fun)), "min_size": int(L[0]),
            "degrees_of_freedom": len(L) - 2}
