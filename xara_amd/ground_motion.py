"""Seeded synthetic ground motion: Kanai-Tajimi filtered Gaussian white noise
with a Jennings envelope, scaled to a target peak ground acceleration.

The white noise comes from numpy's Philox bit generator keyed by the seed
(the frame index in the ensemble), so every motion is reproducible on any
host.  The filter (omega_g = 15.6 rad/s, zeta_g = 0.6) is integrated with
the exact zero-order-hold discretisation of its state equations.
"""
import numpy as np

G_MM_S2 = 9810.0  # standard gravity in mm/s^2 (models are in N, mm)


def _kt_filter(dt, wg, zg):
    A = np.array([[0.0, 1.0], [-wg * wg, -2.0 * zg * wg]])
    lam, V = np.linalg.eig(A)
    Ad = (V @ np.diag(np.exp(lam * dt)) @ np.linalg.inv(V)).real
    Bd = (np.linalg.solve(A, Ad - np.eye(2)) @ np.array([0.0, -1.0])).real
    return Ad, Bd


def kanai_tajimi_batch(seeds, nsteps=2000, dt=0.01, pga_g=0.3, wg=15.6, zg=0.6, t1=2.0, t2=10.0, c=0.25):
    """Motions of several seeds at once, shape (len(seeds), nsteps); row i is
    what ``kanai_tajimi(seeds[i], ..., pga_g[i])`` returns."""
    seeds = np.atleast_1d(seeds)
    pga = np.broadcast_to(np.asarray(pga_g, np.float64), seeds.shape)
    w = np.stack([np.random.Generator(np.random.Philox(int(s))).standard_normal(nsteps) for s in seeds])
    w /= np.sqrt(dt)
    # state x = [xf, vf]; xf'' + 2 zg wg xf' + wg^2 xf = -w ; ag = -(2 zg wg vf + wg^2 xf)
    Ad, Bd = _kt_filter(dt, wg, zg)
    x0 = np.zeros(len(seeds))
    x1 = np.zeros(len(seeds))
    ag = np.zeros((len(seeds), nsteps))
    for k in range(nsteps):
        x0, x1 = Ad[0, 0] * x0 + Ad[0, 1] * x1 + Bd[0] * w[:, k], Ad[1, 0] * x0 + Ad[1, 1] * x1 + Bd[1] * w[:, k]
        ag[:, k] = -(2.0 * zg * wg * x1 + wg * wg * x0)
    t = np.arange(nsteps) * dt
    env = np.where(t < t1, (t / t1) ** 2, np.where(t <= t2, 1.0, np.exp(-c * (t - t2))))
    ag *= env
    ag *= (pga * G_MM_S2 / np.max(np.abs(ag), axis=1))[:, None]
    return ag


def kanai_tajimi(seed, nsteps=2000, dt=0.01, pga_g=0.3, wg=15.6, zg=0.6, t1=2.0, t2=10.0, c=0.25):
    return kanai_tajimi_batch([seed], nsteps, dt, pga_g, wg, zg, t1, t2, c)[0]


def ensemble_pga(seed, n):
    """PGA of each ensemble member, uniform in [0.1, 0.6] g (seeded)."""
    rng = np.random.Generator(np.random.Philox(seed))
    return rng.uniform(0.1, 0.6, n)
