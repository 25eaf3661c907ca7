"""Synthetic groundwater workloads (host-side input generation, not timed).

Each builder returns a ``Problem`` holding exactly the arguments of
``groundwater_steadystate`` (src/groundwater.jl:75) plus the per-cell
``logK`` the edge conductivities come from, the grid geometry and the
observation design.  Index arrays are 1-based like the reference's.

Workloads:

* ``steady_100k``  -- 100x100x10 cells over 50x50x5 m, logK = -9 + GRF
  (Matérn nu=1, correlation length 50 m, sigma 1, seed 0) copied to the 10
  layers, Q = 1e-4 m^3/s at cell 1, Dirichlet h = 0 at the last cell, 50
  observation cells (seed 1); objective sum(h_obs^2).  The problem set-up of
  examples/gw_steadystate_3d/inputdeck.jl:14-24 with harmonic-mean K_e.
* ``steady_2m``    -- 200x200x50 cells over 100x100x25 m, correlation length
  50 m, 4 injectors of 1e-4 m^3/s at seeded free cells (seed 2), Dirichlet
  h = 0 on the top face (i3 = 50).  200 observation cells (seed 3); the
  observed heads come from a second field (seed 4) plus N(0, (1e-3)^2)
  noise (seed 5).
* ``ensemble_members`` -- the 96 realizations of the batched ensemble on
  0.5 m grids.
* ``transient_schedule`` -- the injection ramp of the transient run.
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import fields, grid


def harmonic_Ke(logK, neighbors):
    """K_e = hm(exp(logK_a), exp(logK_b)), hm(x, y) = 2/(1/x + 1/y)
    (src/richards.jl:50-52); host helper for input preparation."""
    k = np.exp(logK)
    ka = k[neighbors[:, 0] - 1]
    kb = k[neighbors[:, 1] - 1]
    return 2.0 / (1.0 / ka + 1.0 / kb)


@dataclass
class Problem:
    name: str
    ns: tuple
    mins: tuple
    maxs: tuple
    neighbors: np.ndarray          # (E, 2) int64, 1-based
    areasoverlengths: np.ndarray   # (E,)
    volumes: np.ndarray            # (n,)
    dirichletnodes: np.ndarray     # (nD,) int64, 1-based
    dirichleths: np.ndarray        # (n,)
    Qs: np.ndarray                 # (n,)
    logK: np.ndarray               # (n,)
    obs: np.ndarray                # observation cells, 1-based
    h_obs: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(self.Qs.shape[0])

    @property
    def nedges(self):
        return int(self.neighbors.shape[0])

    def Ks(self):
        return harmonic_Ke(self.logK, self.neighbors)

    def free_cells(self):
        m = np.ones(self.n, dtype=bool)
        m[self.dirichletnodes - 1] = False
        return np.nonzero(m)[0] + 1


def _seeded_choice(seed, population, k):
    rng = np.random.Generator(np.random.Philox(int(seed)))
    return np.sort(rng.choice(population, size=int(k), replace=False)).astype(np.int64)


def box_problem(name, ns, mins, maxs, lam, sigma, field_seed, mu=-9.0,
                dirichlet="last", sources=((1, 1e-4),), n_injectors=0,
                injector_seed=2, injector_rate=1e-4, n_obs=50, obs_seed=1):
    """A regulargrid3d box with a layered log-conductivity field."""
    coords, nb, aol, vol = grid.regulargrid3d(mins, maxs, ns)
    n = coords.shape[1]
    sp = grid.grid_spacing(mins, maxs, ns)
    logK = fields.layered_logK(ns, sp, mu, lam, sigma, field_seed)
    if dirichlet == "last":
        dn = np.array([n], dtype=np.int64)
    elif dirichlet == "top":
        n3 = int(ns[2])
        dn = np.arange(n3, n + 1, n3, dtype=np.int64)  # i3 == n3 (1-based)
    else:
        raise ValueError(dirichlet)
    dh = np.zeros(n)
    Q = np.zeros(n)
    free = np.ones(n, dtype=bool)
    free[dn - 1] = False
    freecells = np.nonzero(free)[0] + 1
    for cell, rate in sources:
        Q[cell - 1] = rate
    if n_injectors:
        inj = _seeded_choice(injector_seed, freecells, n_injectors)
        Q[inj - 1] = injector_rate
    obs = _seeded_choice(obs_seed, freecells, min(n_obs, freecells.size // 2))
    return Problem(name, tuple(ns), tuple(mins), tuple(maxs), nb, aol, vol, dn, dh,
                   Q, logK, obs, meta=dict(lam=lam, sigma=sigma, field_seed=field_seed))


def steady_100k(sigma=1.0):
    return box_problem("steady_100k", (100, 100, 10), (0.0, 0.0, 0.0), (50.0, 50.0, 5.0),
                       lam=50.0, sigma=sigma, field_seed=0, dirichlet="last",
                       sources=((1, 1e-4),), n_obs=50, obs_seed=1)


def steady_2m(sigma=1.0, ns=(200, 200, 50), maxs=(100.0, 100.0, 25.0)):
    """The 2M-cell problem.  ``ns``/``maxs`` may be shrunk for tests (same
    construction, same seeds)."""
    return box_problem("steady_2m", ns, (0.0, 0.0, 0.0), maxs, lam=50.0, sigma=sigma,
                       field_seed=0, dirichlet="top", sources=(), n_injectors=4,
                       injector_seed=2, n_obs=200, obs_seed=3)


def observation_logK(prob, seed=4):
    """logK of the second field the observed heads are made from."""
    sp = grid.grid_spacing(prob.mins, prob.maxs, prob.ns)
    return fields.layered_logK(prob.ns, sp, -9.0, prob.meta["lam"], prob.meta["sigma"], seed)


def observation_noise(nobs, seed=5, std=1e-3):
    rng = np.random.Generator(np.random.Philox(int(seed)))
    return std * rng.standard_normal(int(nobs))


def fd_cells(prob, k=32, seed=6):
    """Cells whose logK is perturbed in the central finite-difference check."""
    return _seeded_choice(seed, np.arange(1, prob.n + 1), k)


def transient_schedule(nsteps=200, dt=3600.0, ramp_steps=50):
    """Injection ramp Q(t) = Q0 min(1, t / (ramp_steps dt)) at t_n = n dt,
    returned as the per-step scale factor (entry 0 is t = 0)."""
    t = np.arange(nsteps + 1) * dt
    return np.minimum(1.0, t / (ramp_steps * dt))


def ensemble_members(shapes=None):
    """The batched ensemble: 96 realizations, 32 each of 64^3, 96x96x24 and
    128x128x16 cells at 0.5 m node spacing, logK = -9 + layered Matern(nu=1)
    field with sigma 1.5 and a correlation length drawn from U(20, 80) m
    (seeds 100..195), top-face Dirichlet h = 0, 4 injectors of 1e-4 m^3/s and
    50 observation cells per member (the 2M-cell pattern scaled).  ``shapes``
    (optional) replaces the list of member grid shapes; member m keeps seed
    100 + m.  Returns dicts with n, nb, aol, dn, logK, Q, obs, ns, lam."""
    if shapes is None:
        shapes = [(64, 64, 64)] * 32 + [(96, 96, 24)] * 32 + [(128, 128, 16)] * 32
    grids = {}
    out = []
    for m, ns in enumerate(shapes):
        ns = tuple(ns)
        seed = 100 + m
        lam = float(np.random.Generator(np.random.Philox(seed)).uniform(20.0, 80.0))
        maxs = tuple(0.5 * (k - 1) for k in ns)
        if ns not in grids:
            grids[ns] = grid.regulargrid3d((0.0, 0.0, 0.0), maxs, ns)
        coords, nb, aol, vol = grids[ns]
        n = coords.shape[1]
        logK = fields.layered_logK(ns, (0.5, 0.5, 0.5), -9.0, lam, 1.5, seed)
        dn = np.arange(ns[2], n + 1, ns[2], dtype=np.int64)
        free = np.ones(n, dtype=bool)
        free[dn - 1] = False
        fc = np.nonzero(free)[0] + 1
        rng = np.random.Generator(np.random.Philox(seed + 1000))
        inj = rng.choice(fc, 4, replace=False)
        obs = np.sort(rng.choice(fc, 50, replace=False))
        Q = np.zeros(n)
        Q[inj - 1] = 1e-4
        out.append(dict(n=n, nb=nb, aol=aol, dn=dn, logK=logK, Q=Q, obs=obs, ns=ns, lam=lam))
    return out
