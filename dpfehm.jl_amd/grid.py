"""Structured-grid builders with DPFEHM's cell and edge ordering.

Host-side input generation (not timed, not on the GPU).  Mirrors
``src/grid.jl`` of DPFEHM.jl:

* ``getpoints``      -- src/grid.jl:27-29
* ``regulargrid1d``  -- src/grid.jl:79-96
* ``regulargrid2d``  -- src/grid.jl:109-138
* ``regulargrid3d``  -- src/grid.jl:150-188
* ``load_uge``       -- src/grid.jl:38-65 (PFLOTRAN explicit unstructured grid)

Conventions kept from the reference:

* node-centred spacing ``dx = (max - min) / (n - 1)``;
* 3-D linear index ``i3 + n3*(i2-1) + n3*n2*(i1-1)`` (z fastest, then y,
  then x), 2-D index ``i2 + n2*(i1-1)``;
* edges are emitted per cell in linear order, x-edge first, then y, then z;
* ``neighbors`` are 1-based ``(a, b)`` pairs (Julia ``Pair{Int,Int}``),
  returned as an ``(E, 2)`` int64 array whose rows are the pairs;
* ``coords`` is ``(dim, n)`` like the reference's column-per-cell matrix.
"""
import numpy as np


def getpoints(lo, hi, n):
    """``range(min, max; length=n)`` (src/grid.jl:27-29)."""
    return np.linspace(float(lo), float(hi), int(n))


def regulargrid1d(mins, maxs, ns, dy, dz):
    """src/grid.jl:79-96.  As in the reference, ``areasoverlengths`` has
    ``ns`` entries although there are ``ns - 1`` edges (the last is unused)."""
    n1 = int(ns[0] if np.ndim(ns) else ns)
    lo = mins[0] if np.ndim(mins) else mins
    hi = maxs[0] if np.ndim(maxs) else maxs
    xs = getpoints(lo, hi, n1)
    dx = xs[1] - xs[0]
    coords = xs[None, :].copy()
    idx = np.arange(1, n1, dtype=np.int64)
    neighbors = np.stack([idx, idx + 1], axis=1)
    areasoverlengths = np.full(n1, dy * dz / dx)
    volumes = np.full(n1, dx * dy * dz)
    return coords, neighbors, areasoverlengths, volumes


def regulargrid2d(mins, maxs, ns, dz):
    """src/grid.jl:109-138."""
    n1, n2 = int(ns[0]), int(ns[1])
    xs = getpoints(mins[0], maxs[0], n1)
    ys = getpoints(mins[1], maxs[1], n2)
    dx = xs[1] - xs[0]
    dy = ys[1] - ys[0]
    i1, i2 = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    lin = (i2 + n2 * i1).ravel()  # 0-based, increasing in loop order
    coords = np.empty((2, n1 * n2))
    coords[0, lin] = xs[i1.ravel()]
    coords[1, lin] = ys[i2.ravel()]
    cand_b = np.stack([lin + n2, lin + 1], axis=1)
    valid = np.stack([(i1 < n1 - 1).ravel(), (i2 < n2 - 1).ravel()], axis=1)
    aol_c = np.broadcast_to(np.array([dy * dz / dx, dx * dz / dy]), cand_b.shape)
    a = np.broadcast_to(lin[:, None], cand_b.shape)
    neighbors = np.stack([a[valid], cand_b[valid]], axis=1).astype(np.int64) + 1
    areasoverlengths = aol_c[valid].copy()
    volumes = np.full(n1 * n2, dx * dy * dz)
    return coords, neighbors, areasoverlengths, volumes


def regulargrid3d(mins, maxs, ns):
    """src/grid.jl:150-188."""
    n1, n2, n3 = int(ns[0]), int(ns[1]), int(ns[2])
    xs = getpoints(mins[0], maxs[0], n1)
    ys = getpoints(mins[1], maxs[1], n2)
    zs = getpoints(mins[2], maxs[2], n3)
    dx = xs[1] - xs[0]
    dy = ys[1] - ys[0]
    dz = zs[1] - zs[0]
    n = n1 * n2 * n3
    lin = np.arange(n, dtype=np.int64)  # loop order i1, i2, i3 == linear order
    i1 = lin // (n2 * n3)
    i2 = (lin // n3) % n2
    i3 = lin % n3
    coords = np.empty((3, n))
    coords[0] = xs[i1]
    coords[1] = ys[i2]
    coords[2] = zs[i3]
    cand_b = np.stack([lin + n3 * n2, lin + n3, lin + 1], axis=1)
    valid = np.stack([i1 < n1 - 1, i2 < n2 - 1, i3 < n3 - 1], axis=1)
    aol_c = np.broadcast_to(
        np.array([dy * dz / dx, dx * dz / dy, dx * dy / dz]), cand_b.shape)
    a = np.broadcast_to(lin[:, None], cand_b.shape)
    neighbors = np.stack([a[valid], cand_b[valid]], axis=1) + 1
    areasoverlengths = aol_c[valid].copy()
    volumes = np.full(n, dx * dy * dz)
    return coords, neighbors, areasoverlengths, volumes


def grid_spacing(mins, maxs, ns):
    """Node spacing per axis, ``(max - min) / (n - 1)`` (src/grid.jl:156-158)."""
    return tuple((float(b) - float(a)) / (int(k) - 1) for a, b, k in zip(mins, maxs, ns))


def load_uge(filename):
    """src/grid.jl:38-65: read an ASCII PFLOTRAN explicit unstructured grid
    (``.uge``) and return ``coords`` (3, n), ``volumes`` (n), ``neighbors``
    (E, 2) 1-based, ``areas`` (E) and ``lengths`` (E), the distance between
    the two cell centres of each connection.

    The file holds ``CELLS n`` followed by n lines ``id x y z volume``, then
    ``CONNECTIONS m`` followed by m lines ``id1 id2 x y z area``.  As in the
    reference, the cell ids of the CELLS block are not read (cells are taken
    in file order), the face centre (x, y, z) of a connection is not used,
    and a malformed line raises.  ``areas / lengths`` is the
    ``areasoverlengths`` argument of the solvers."""
    with open(filename) as f:
        def line():
            s = f.readline()
            if not s:
                raise ValueError(f"{filename}: unexpected end of file")
            return s.split()
        numcells = int(line()[1])
        coords = np.empty((3, numcells))
        volumes = np.empty(numcells)
        for i in range(numcells):
            parts = line()
            coords[:, i] = [float(parts[1]), float(parts[2]), float(parts[3])]
            volumes[i] = float(parts[4])
        numconnections = int(line()[1])
        neighbors = np.empty((numconnections, 2), dtype=np.int64)
        areas = np.empty(numconnections)
        for i in range(numconnections):
            parts = line()
            neighbors[i] = int(parts[0]), int(parts[1])
            areas[i] = float(parts[5])
    d = coords[:, neighbors[:, 0] - 1] - coords[:, neighbors[:, 1] - 1]
    lengths = np.sqrt((d * d).sum(axis=0))
    return coords, volumes, neighbors, areas, lengths
