"""Peng-Robinson fugacity (mcmc/fugacity.py:36-193), host side (numpy).

Solves the PR cubic for the compressibility factor Z of a gas mixture and
returns ln f_i = ln phi_i + ln y_i + ln P, choosing the real positive root of
minimum ln phi (the gas phase, as the reference does after RASPA).
"""

from __future__ import annotations

import numpy as np


def peng_robinson_log_fugacity(pressure, temperature, critical_pressure, critical_temperature,
                               acentric_factor, fraction=None, interaction=None):
    """ln fugacity [ln(pressure unit)] per component and Z of the chosen root."""
    Pc = np.atleast_1d(np.asarray(critical_pressure, float))
    Tc = np.atleast_1d(np.asarray(critical_temperature, float))
    w = np.atleast_1d(np.asarray(acentric_factor, float))
    n = Pc.size
    y = np.ones(n) / n if fraction is None else np.asarray(fraction, float)
    kij = np.zeros((n, n)) if interaction is None else np.asarray(interaction, float)
    Pr = pressure / Pc
    Tr = temperature / Tc
    m = 0.37464 + 1.54226 * w - 0.26992 * w**2
    sqrt_alpha = 1 + m * (1 - Tr**0.5)
    A = 0.45724 * Pr / Tr**2 * sqrt_alpha**2
    B = 0.0778 * Pr / Tr
    Aij = (1 - kij) * np.sqrt(np.outer(A, A))
    Bm = y @ B
    Am = y @ Aij @ y
    roots = np.roots([1.0, Bm - 1, Am - 3.0 * Bm**2 - 2 * Bm, -(Am * Bm - Bm**2 - Bm**3)])
    Z = np.where(np.abs(roots.imag) < 1e-12, roots.real, np.nan)
    sq2 = np.sqrt(2.0)
    dA = 2 * (Aij @ y)[:, None]
    with np.errstate(invalid="ignore", divide="ignore"):
        ln_phi = ((B[:, None] / Bm) * (Z - 1) - np.log(Z - Bm)
                  - (Am / (2 * sq2 * Bm)) * (dA / Am - B[:, None] / Bm)
                  * np.log((Z + (1 + sq2) * Bm) / (Z + (1 - sq2) * Bm)))
    valid = np.isfinite(Z) & (Z > 0)
    ln_phi_v = np.where(valid[None, :] & np.isfinite(ln_phi), ln_phi, np.inf)
    idx = np.argmin(ln_phi_v, axis=1)
    ln_f = ln_phi[np.arange(n), idx] + np.log(pressure) + np.log(y)
    return ln_f, Z[idx]
