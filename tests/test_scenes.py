"""Synthetic scenes: geometry of the bodies and the rigid-motion stepping."""
import numpy as np

from fastmbcui_amd import scenes


def test_rounded_box_points_lie_on_the_grown_box_with_normals():
    rng = np.random.default_rng(0)
    half = np.array([1.0, 0.7, 0.5])
    p, n = scenes._rounded_box_surface(rng, 20000, half, 0.1)
    inner = half - 0.1
    q = np.clip(p, -inner, inner)            # closest point of the inner box
    d = np.linalg.norm(p - q, axis=1)
    assert np.allclose(d, 0.1, atol=1e-12)
    assert np.allclose((p - q) / d[:, None], n, atol=1e-12)


def test_rounded_box_sampling_is_area_uniform():
    # the fraction of points on the flat +z face matches its share of the area
    rng = np.random.default_rng(1)
    half = np.array([1.2, 0.9, 0.6])
    r = 0.1
    p, n = scenes._rounded_box_surface(rng, 200000, half, r)
    e = 2 * (half - r)
    faces = 2 * (e[1] * e[2] + e[0] * e[2] + e[0] * e[1])
    edges = sum(4 * e[d] * np.pi * r / 2 for d in range(3))
    total = faces + edges + 4 * np.pi * r * r
    share = np.mean(n[:, 2] > 1 - 1e-12)
    assert abs(share - e[0] * e[1] / total) < 0.005


def test_rigid_motion_preserves_body_geometry_and_matches_rhs():
    rs = scenes.mixed_bodies(seed=2, n_bodies=6, pts_per_body=64, box=12.0)
    s0 = rs.scene()
    rs.advance(0.05)
    s1 = rs.scene()
    for bi in range(6):
        sel = s0.body_of_point == bi
        d0 = np.linalg.norm(s0.x[sel][:, None] - s0.x[sel][None], axis=2)
        d1 = np.linalg.norm(s1.x[sel][:, None] - s1.x[sel][None], axis=2)
        assert np.allclose(d0, d1, atol=1e-12)
        # sources keep their depth below the surface along the normal
        assert np.allclose(np.einsum("ij,ij->i", s1.x[sel] - s1.y[sel], s1.nx[sel]),
                           np.einsum("ij,ij->i", s0.x[sel] - s0.y[sel], s0.nx[sel]), atol=1e-12)
    # b = n . (v + w x (x - c)) is the rate of change of the points along n
    h = 1e-6
    rs2 = scenes.mixed_bodies(seed=2, n_bodies=6, pts_per_body=64, box=12.0)
    a = rs2.scene()
    rs2.advance(h)
    bpt = rs2.scene()
    fd = np.einsum("ij,ij->i", (bpt.x - a.x) / h, a.nx)
    assert np.allclose(fd, a.b[:, 0], atol=1e-5)


def test_mixed_bodies_shapes_and_sizes():
    rs = scenes.mixed_bodies(seed=0, n_bodies=9, pts_per_body=128, box=14.0)
    sc = rs.scene()
    assert sc.n == sc.m == 9 * 128
    assert list(rs.meta["shapes"][:3]) == [0, 1, 2]
    assert np.allclose(np.linalg.norm(sc.nx, axis=1), 1.0)
