"""Seeded log-conductivity fields (host-side input generation, not timed).

The reference examples draw ``logK`` from GaussianRandomFields.jl with a
Matérn covariance (``examples/gw_steadystate_3d/inputdeck.jl:27-41``) and
copy the 2-D field to every z-layer.  That library is not available here, so
fields are made by spectral synthesis of the same covariance family:

* Matérn with smoothness ``nu`` and correlation length ``lam`` in the
  GaussianRandomFields convention ``C(r) = s^2 2^(1-nu)/Gamma(nu)
  (sqrt(2 nu) r/lam)^nu K_nu(sqrt(2 nu) r/lam)``, whose 2-D spectral density
  is proportional to ``(2 nu / lam^2 + |k|^2)^-(nu + 1)``;
* white noise from numpy's Philox bit generator keyed by ``seed``;
* a periodic embedding padded to twice the grid in each direction, cropped
  back, with the spectrum normalised so that the point variance is
  ``sigma^2``.

The same seed always gives the same field bit for bit on this numpy, and every
path (GPU, oracle, CPU baseline) is fed the identical array.
"""
import numpy as np


def matern_field_2d(nx, ny, dx, dy, lam, sigma, seed, nu=1.0):
    """Gaussian random field on an ``nx`` x ``ny`` node lattice, indexed
    ``[ix, iy]``."""
    px, py = 2 * int(nx), 2 * int(ny)
    rng = np.random.Generator(np.random.Philox(int(seed)))
    xi = rng.standard_normal((px, py))
    kx = 2.0 * np.pi * np.fft.fftfreq(px, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(py, d=dy)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    spec = (2.0 * nu / lam ** 2 + k2) ** (-(nu + 1.0))
    spec *= sigma ** 2 / spec.mean()
    g = np.fft.ifft2(np.sqrt(spec) * np.fft.fft2(xi, norm="ortho"), norm="ortho").real
    return g[:nx, :ny].copy()


def layered_logK(ns, spacing, mu, lam, sigma, seed):
    """``mu + GRF`` on the (x, y) nodes, copied to every z-layer, returned in
    the reference's linear cell order (z fastest), as in
    ``examples/gw_steadystate_3d/inputdeck.jl:36-41``."""
    n1, n2, n3 = (int(v) for v in ns)
    f = matern_field_2d(n1, n2, spacing[0], spacing[1], lam, sigma, seed)
    return (mu + np.repeat(f.reshape(n1 * n2), n3)).astype(np.float64)
