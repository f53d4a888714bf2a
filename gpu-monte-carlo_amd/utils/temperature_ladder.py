"""Temperature ladders built in inverse temperature
(utils/temperature_ladder.py of the reference).

``Tc=None`` gives a ladder uniform in beta.  Otherwise the density of betas is
raised inside a window of full width ``tc_window`` around ``Tc`` (smooth tanh
shoulders) so that a fraction ``tc_fraction`` of the ladder falls inside it.
Temperatures are returned in ascending order.
"""

import numpy as np


def make_temperature_ladder(T_min, T_max, n_T, Tc=None, tc_window=0.08, tc_fraction=0.60):
    if not (0 < T_min < T_max):
        raise ValueError("Require 0 < T_min < T_max")
    if n_T < 2:
        raise ValueError("Need at least two temperatures")
    beta_min, beta_max = 1.0 / T_max, 1.0 / T_min
    if Tc is None:
        return (1.0 / np.linspace(beta_min, beta_max, n_T))[::-1]

    T_lo, T_hi = Tc - tc_window / 2, Tc + tc_window / 2
    if not T_min < T_lo < T_hi < T_max:
        raise ValueError("Focus window must lie inside range")
    if not 0 < tc_fraction < 1:
        raise ValueError("tc_fraction must be between 0 and 1")
    b_lo, b_hi = 1.0 / T_hi, 1.0 / T_lo

    beta = np.linspace(beta_min, beta_max, 4097)
    shoulder = (b_hi - b_lo) / 8
    profile = 0.5 * (np.tanh((beta - b_lo) / shoulder) - np.tanh((beta - b_hi) / shoulder))

    # amplitude A of density 1 + A * profile that puts tc_fraction inside the window
    span, width = beta_max - beta_min, b_hi - b_lo
    total = np.trapezoid(profile, beta)
    inner = np.r_[b_lo, beta[(beta > b_lo) & (beta < b_hi)], b_hi]
    inside = np.trapezoid(np.interp(inner, beta, profile), inner)
    numerator = tc_fraction * span - width
    denominator = inside - tc_fraction * total
    if numerator < -1e-12 or denominator <= 0:
        raise ValueError("Requested focus fraction is not achievable")
    density = 1.0 + max(0.0, numerator / denominator) * profile

    cdf = np.r_[0.0, np.cumsum(0.5 * (density[1:] + density[:-1]) * np.diff(beta))]
    cdf /= cdf[-1]
    betas = np.interp(np.linspace(0, 1, n_T), cdf, beta)
    return (1.0 / betas)[::-1]
