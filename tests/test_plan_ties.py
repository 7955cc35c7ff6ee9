"""CPU checks of the product's host maximin ordering (mfs_order_maximin, the
plan's first step) against the oracle on tie-heavy inputs: lattices, rings
and an unjittered Fibonacci sphere, where many candidate distances are
exactly equal and the heap's tie rules decide the order
(src/maxmin.h:194-293, 588-665; root rule src/preprocess.h:785-823).

Both sides follow the reference's rules line by line; these inputs make every
tie rule fire, so a difference in any of them shows as a different order.
No device work: the ordering is host code in libmfs_core.so."""
import numpy as np
import pytest

import oracle
from fastmbcui_amd import _lib


def lattice(nx, ny, nz, h=1.0):
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    return (g.reshape(-1, 3) * h).astype(np.float64)


def rings(count, per_ring, radius=1.0, gap=0.5):
    pts = []
    for r in range(count):
        t = 2 * np.pi * np.arange(per_ring) / per_ring
        pts.append(np.stack([radius * np.cos(t), radius * np.sin(t), np.full(per_ring, r * gap)], -1))
    return np.concatenate(pts)


def fibonacci_sphere(npts, radius=1.0):
    k = np.arange(npts)
    z = 1.0 - (2.0 * k + 1.0) / npts
    r = np.sqrt(1.0 - z * z)
    phi = k * np.pi * (3.0 - np.sqrt(5.0))
    return radius * np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)


TIE_CASES = {
    "lattice_8x8x8": lambda: lattice(8, 8, 8),
    "lattice_17x5x3": lambda: lattice(17, 5, 3, 0.25),
    "plane_grid_20x20": lambda: lattice(20, 20, 1),
    "rings_6x48": lambda: rings(6, 48),
    "rings_3x7": lambda: rings(3, 7, 2.0, 1.0),
    "fibonacci_1024": lambda: fibonacci_sphere(1024),
    "line_100": lambda: lattice(100, 1, 1),
}


@pytest.mark.parametrize("name", sorted(TIE_CASES))
def test_host_ordering_equals_oracle_on_ties(name):
    pts = TIE_CASES[name]()
    P, Pinv, scale = _lib.order_maximin(pts)
    oP, oPinv, oscale, _, _ = oracle.fast_fps(pts)
    assert np.array_equal(P, oP), name
    assert np.array_equal(Pinv, oPinv), name
    assert np.array_equal(scale, oscale), name


@pytest.mark.parametrize("name", ["lattice_8x8x8", "rings_6x48", "fibonacci_1024"])
def test_host_ordering_distances_are_maximin_on_ties(name):
    """Whatever the tie order, each pivot's length scale is its exact distance
    to the earlier (coarser) pivots, and the scales ascend."""
    pts = TIE_CASES[name]()
    P, Pinv, scale = _lib.order_maximin(pts)
    n = pts.shape[0]
    assert np.all(np.diff(scale) >= 0)
    assert np.array_equal(np.sort(Pinv), np.arange(n))
    # new index n-1 is the root; pivot new index c has earlier pivots c+1 .. n-1
    q = pts[Pinv]
    for c in range(0, n - 1, max(1, n // 97)):
        d = np.sqrt(((q[c + 1:] - q[c]) ** 2).sum(1)).min()
        assert d == pytest.approx(scale[c], rel=0, abs=1e-12)


def test_host_ordering_rejects_coincident_points():
    pts = lattice(4, 4, 4)
    pts[10] = pts[3]
    with pytest.raises(_lib.MfsError) as e:
        _lib.order_maximin(pts)
    assert e.value.code == _lib.MFS_ERR_DUPLICATE
