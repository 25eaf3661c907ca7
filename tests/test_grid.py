"""Grid builders: the layout invariants of the reference's test/grid.jl:9-24
plus edge counts and order (src/grid.jl:150-188)."""
import numpy as np


def test_grid2d_layout(pkg):
    ns, mins, maxs = [5, 10, 20], [0, 0, 0], [1, 1, 1]
    coords, nb, aol, vol = pkg.grid.regulargrid2d(mins[:2], maxs[:2], ns[:2], maxs[2] - mins[2])
    assert coords[1, 0] != coords[1, 1]
    assert coords[1, 0] == coords[1, ns[1]]
    assert coords[0, 0] != coords[0, ns[1]]
    assert np.all(coords[0, 0] == coords[0, :ns[1]])
    expect = np.prod(ns[:2]) * np.prod((np.array(maxs[:2]) - mins[:2]) / (np.array(ns[:2]) - 1))
    assert np.isclose(vol.sum(), expect * (maxs[2] - mins[2]))
    assert nb.shape == (2 * ns[0] * ns[1] - ns[0] - ns[1], 2)


def test_grid3d_layout(pkg):
    ns, mins, maxs = [5, 10, 20], [0, 0, 0], [1, 1, 1]
    coords, nb, aol, vol = pkg.grid.regulargrid3d(mins, maxs, ns)
    assert coords[2, 0] != coords[2, 1]
    assert coords[2, 0] == coords[2, ns[2]]
    assert coords[1, 0] != coords[1, ns[2]]
    assert np.all(coords[1, 0] == coords[1, :ns[2]])
    assert coords[0, 0] != coords[0, ns[2] * ns[1]]
    assert np.all(coords[0, 0] == coords[0, :ns[1] * ns[2]])
    expect = np.prod(ns) * np.prod((np.array(maxs) - mins) / (np.array(ns) - 1))
    assert np.isclose(vol.sum(), expect)
    n1, n2, n3 = ns
    E = 3 * n1 * n2 * n3 - n1 * n2 - n1 * n3 - n2 * n3
    assert nb.shape == (E, 2) and aol.shape == (E,)


def test_grid3d_edge_order_matches_loop(pkg):
    """Brute-force restatement of the triple loop of src/grid.jl:164-186."""
    ns = (3, 4, 5)
    mins, maxs = (0, 0, 0), (2.0, 3.0, 8.0)
    _, nb, aol, _ = pkg.grid.regulargrid3d(mins, maxs, ns)
    dx, dy, dz = 1.0, 1.0, 2.0
    lin = lambda i1, i2, i3: i3 + ns[2] * (i2 - 1) + ns[2] * ns[1] * (i1 - 1)
    ref_nb, ref_aol = [], []
    for i1 in range(1, ns[0] + 1):
        for i2 in range(1, ns[1] + 1):
            for i3 in range(1, ns[2] + 1):
                if i1 < ns[0]:
                    ref_nb.append((lin(i1, i2, i3), lin(i1 + 1, i2, i3))); ref_aol.append(dy * dz / dx)
                if i2 < ns[1]:
                    ref_nb.append((lin(i1, i2, i3), lin(i1, i2 + 1, i3))); ref_aol.append(dx * dz / dy)
                if i3 < ns[2]:
                    ref_nb.append((lin(i1, i2, i3), lin(i1, i2, i3 + 1))); ref_aol.append(dx * dy / dz)
    assert np.array_equal(nb, np.array(ref_nb))
    assert np.allclose(aol, ref_aol, rtol=0, atol=1e-15)


def test_grid1d_matches_reference_lengths(pkg):
    coords, nb, aol, vol = pkg.grid.regulargrid1d([0.0], [10.0], [11], 2.0, 3.0)
    assert nb.shape == (10, 2) and aol.shape == (11,)  # reference quirk: length ns
    assert np.allclose(aol, 6.0) and np.allclose(vol, 6.0)


def test_fields_seeded_and_scaled(pkg):
    a = pkg.fields.matern_field_2d(100, 100, 0.5, 0.5, 50.0, 1.0, seed=0)
    b = pkg.fields.matern_field_2d(100, 100, 0.5, 0.5, 50.0, 1.0, seed=0)
    c = pkg.fields.matern_field_2d(100, 100, 0.5, 0.5, 50.0, 1.0, seed=1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # point variance sigma^2 over an ensemble of seeds
    v = np.mean([pkg.fields.matern_field_2d(32, 32, 4.0, 4.0, 10.0, 2.0, seed=s)[5, 7] ** 2
                 for s in range(400)])
    assert 2.6 < v < 5.6


def test_workload_sizes(pkg):
    w = pkg.workloads.steady_2m(ns=(20, 20, 5), maxs=(10.0, 10.0, 2.5))
    assert w.n == 2000 and w.dirichletnodes.size == 400
    assert np.all(w.dirichletnodes % 5 == 0)
    assert np.count_nonzero(w.Qs) == 4 and np.all(w.Qs[w.Qs != 0] == 1e-4)
    assert np.intersect1d(np.nonzero(w.Qs)[0] + 1, w.dirichletnodes).size == 0
    assert w.obs.size == 200 and np.intersect1d(w.obs, w.dirichletnodes).size == 0


def _write_uge(path, coords, volumes, nb, areas, shuffle_ids=False):
    with open(path, "w") as f:
        f.write(f"CELLS {coords.shape[1]}\n")
        for i in range(coords.shape[1]):
            cid = coords.shape[1] - i if shuffle_ids else i + 1
            x, y, z, v = (float(t) for t in (*coords[:, i], volumes[i]))
            f.write(f"{cid} {x!r} {y!r} {z!r} {v!r}\n")
        f.write(f"CONNECTIONS {nb.shape[0]}\n")
        for e in range(nb.shape[0]):
            c = [float(t) for t in 0.5 * (coords[:, nb[e, 0] - 1] + coords[:, nb[e, 1] - 1])]
            f.write(f"{nb[e, 0]} {nb[e, 1]} {c[0]!r} {c[1]!r} {c[2]!r} {float(areas[e])!r}\n")


def test_load_uge_reproduces_regular_grid(pkg, tmp_path):
    """A regular grid written as a .uge (face areas dy*dz, dx*dz, dx*dy) reads
    back with the same neighbours and volumes, lengths equal to the grid
    spacing along each edge's axis, and areas / lengths equal to
    regulargrid3d's areasoverlengths (src/grid.jl:38-65 against :150-188)."""
    ns, mins, maxs = (4, 3, 5), (0.0, 0.0, 0.0), (3.0, 4.0, 2.0)
    coords, nb, aol, vol = pkg.grid.regulargrid3d(mins, maxs, ns)
    dx, dy, dz = pkg.grid.grid_spacing(mins, maxs, ns)
    step = nb[:, 1] - nb[:, 0]
    spacing = np.where(step == ns[1] * ns[2], dx, np.where(step == ns[2], dy, dz))
    areas = aol * spacing
    path = tmp_path / "box.uge"
    _write_uge(path, coords, vol, nb, areas, shuffle_ids=True)
    c2, v2, nb2, a2, l2 = pkg.grid.load_uge(str(path))
    assert np.array_equal(nb2, nb) and nb2.dtype == np.int64
    assert np.array_equal(c2, coords) and np.array_equal(v2, vol) and np.array_equal(a2, areas)
    assert np.allclose(l2, spacing, rtol=1e-15, atol=0)
    assert np.allclose(a2 / l2, aol, rtol=1e-14, atol=0)


def test_load_uge_truncated_file_raises(pkg, tmp_path):
    path = tmp_path / "bad.uge"
    path.write_text("CELLS 2\n1 0 0 0 1\n")
    import pytest
    with pytest.raises(ValueError):
        pkg.grid.load_uge(str(path))
