"""Seeded synthetic ground motion: Kanai-Tajimi filtered Gaussian white noise
with a Jennings envelope, scaled to a target peak ground acceleration.

The white noise comes from numpy's Philox bit generator keyed by the seed
(the frame index in the ensemble), so every motion is reproducible on any
host.  The filter (omega_g = 15.6 rad/s, zeta_g = 0.6) is integrated with
the exact zero-order-hold discretisation of its state equations.
"""
import numpy as np

G_MM_S2 = 9810.0  # standard gravity in mm/s^2 (models are in N, mm)


def kanai_tajimi(seed, nsteps=2000, dt=0.01, pga_g=0.3, wg=15.6, zg=0.6, t1=2.0, t2=10.0, c=0.25):
    rng = np.random.Generator(np.random.Philox(seed))
    w = rng.standard_normal(nsteps) / np.sqrt(dt)
    # state x = [xf, vf]; xf'' + 2 zg wg xf' + wg^2 xf = -w ; ag = -(2 zg wg vf + wg^2 xf)
    A = np.array([[0.0, 1.0], [-wg * wg, -2.0 * zg * wg]])
    lam, V = np.linalg.eig(A)
    Ad = (V @ np.diag(np.exp(lam * dt)) @ np.linalg.inv(V)).real
    Bd = (np.linalg.solve(A, Ad - np.eye(2)) @ np.array([0.0, -1.0])).real
    x = np.zeros(2)
    ag = np.zeros(nsteps)
    for k in range(nsteps):
        x = Ad @ x + Bd * w[k]
        ag[k] = -(2.0 * zg * wg * x[1] + wg * wg * x[0])
    t = np.arange(nsteps) * dt
    env = np.where(t < t1, (t / t1) ** 2, np.where(t <= t2, 1.0, np.exp(-c * (t - t2))))
    ag *= env
    ag *= pga_g * G_MM_S2 / np.max(np.abs(ag))
    return ag


def ensemble_pga(seed, n):
    """PGA of each ensemble member, uniform in [0.1, 0.6] g (seeded)."""
    rng = np.random.Generator(np.random.Philox(seed))
    return rng.uniform(0.1, 0.6, n)
