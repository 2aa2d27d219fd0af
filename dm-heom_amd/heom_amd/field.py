"""Time-dependent electric fields for heom_propagate_timedep: the reference's e_field
(dm-heom/include/heom/e_field.hpp) as stage Hamiltonians.

The reference updates the Hamiltonian before every RHS evaluation with the field at that
evaluation's time, H(t) = H0 - E(t) D- - conj(E(t)) D+ (e_field.hpp:78), where the
pump/probe field is a sum of two Gaussian pulses with carrier phase and frequency
(e_field.hpp:45-58).  Classical RK4 evaluates at t_n, t_n + h/2, t_n + h/2, t_n + h.
"""
from __future__ import annotations

import numpy as np

RK4_STAGE_TIMES = np.array([0.0, 0.5, 0.5, 1.0])


def gaussian_pulse(efield: float, center: float, width: float, phase: float, frequency: float):
    """E(t) = efield exp(-(t - center)^2 / (2 width^2)) exp(i (phase + frequency t)),
    one term of e_field::compute_field (e_field.hpp:49-56); times in s, frequency in rad/s."""
    def field(t):
        return efield * np.exp(-((t - center) ** 2) / (2.0 * width * width)) * np.exp(1j * (phase + frequency * t))
    return field


def pump_probe(probe: dict, pump: dict):
    """The probe plus the pump pulse (e_field.hpp:45-58); each a dict of gaussian_pulse's
    arguments."""
    fp, fq = gaussian_pulse(**probe), gaussian_pulse(**pump)
    return lambda t: fp(t) + fq(t)


def stage_hamiltonians(H0, d_minus, d_plus, field, t0: float, h: float, nsteps: int) -> np.ndarray:
    """[nsteps, 4, N, N] complex: H0 - E(t) D- - conj(E(t)) D+ at the RK4 stage times of
    steps t0 + n h, n = 0 .. nsteps - 1."""
    H0, dm, dp = (np.asarray(x, dtype=np.complex128) for x in (H0, d_minus, d_plus))
    times = t0 + h * (np.arange(nsteps)[:, None] + RK4_STAGE_TIMES[None, :])
    E = np.asarray(field(times), dtype=np.complex128)
    return H0[None, None] - E[..., None, None] * dm[None, None] - np.conj(E)[..., None, None] * dp[None, None]
