"""Oracle heightfield colliders: a 3D heightfield is the trimesh of its samples
(src/rbd/shapes.rs:252-268), built by slosh_amd.trimesh.heightfield and
projected by the trimesh walk (collision/trimesh.slang). Pinned by a closed form
(a sloped plane is its own triangulation: orthogonal projection, inside below)
and by brute force on a bumpy field (closest distance over all triangles; inside
iff below the triangulated surface, for points whose closest point cannot reach
the border)."""
import numpy as np

from oracle import oracle as O
from slosh_amd import models as M
from slosh_amd import trimesh as TM
from tests.test_oracle_trimesh import closest_on_triangle, pose, quat_rotate


def surface_height(v, t, x, z):
    """Height of the triangulated surface above (x, z) (barycentric in the covering triangle)."""
    for tri in t:
        a, b, c = v[tri].astype(np.float64)
        d = (b[2] - c[2]) * (a[0] - c[0]) + (c[0] - b[0]) * (a[2] - c[2])
        l1 = ((b[2] - c[2]) * (x - c[0]) + (c[0] - b[0]) * (z - c[2])) / d
        l2 = ((c[2] - a[2]) * (x - c[0]) + (a[0] - c[0]) * (z - c[2])) / d
        l3 = 1.0 - l1 - l2
        if min(l1, l2, l3) >= -1e-12:
            return l1 * a[1] + l2 * b[1] + l3 * c[1]
    raise AssertionError("point outside the heightfield footprint")


def project(v, t, po, world, max_dist):
    buf = TM.CollisionBuffers()
    shape = TM.convert(v, t, buf)
    return O.project_point_on_boundary(shape, po, world, (buf.vertex_array(), buf.index_array()), max_dist)


def test_heightfield_layout_and_orientation():
    hgt = np.arange(12, dtype=np.float64).reshape(3, 4)
    v, t = TM.heightfield(hgt, (2.0, 0.5, 3.0))
    assert v.shape == (12, 3) and t.shape == (12, 3)
    # Sample (i, j): columns along x, rows along z, unit square centred at 0, scaled.
    assert np.allclose(v[1 * 4 + 2], [(2 / 3 - 0.5) * 2.0, 6 * 0.5, 0.0])
    n = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    assert np.all(n[:, 1] > 0)
    # The triangles tile the footprint exactly once (areas in the xz plane).
    assert np.isclose(0.5 * np.abs(n[:, 1]).sum(), 2.0 * 3.0)


def test_sloped_plane_closed_form():
    # heights linear in (i, j): every triangulation is the plane y = gx x + gz z.
    nr, nc, scale = 7, 9, (1.2, 0.8, 1.0)
    i, j = np.meshgrid(np.arange(nr), np.arange(nc), indexing="ij")
    hgt = 0.05 * i - 0.03 * j
    v, t = TM.heightfield(hgt, scale)
    gx = -0.03 * scale[1] * (nc - 1) / scale[0]
    gz = 0.05 * scale[1] * (nr - 1) / scale[2]
    y0 = (0.05 * (nr - 1) / 2 - 0.03 * (nc - 1) / 2) * scale[1]
    nrm = np.array([-gx, 1.0, -gz]) / np.linalg.norm([-gx, 1.0, -gz])
    po = pose(M.quat_from_scaled_axis((0.2, -0.4, 0.1)), (0.3, -0.2, 0.1), 1.3)
    q = np.array(po["rotation"], np.float64)
    tr = np.array(po["translation_scale"][:3], np.float64)
    sc = float(po["translation_scale"][3])
    qinv = np.array([-q[0], -q[1], -q[2], q[3]])
    rng = np.random.default_rng(5)
    seen = {True: 0, False: 0}
    for _ in range(400):
        x, z = rng.uniform(-0.3, 0.3, 2) * np.array([scale[0], scale[2]])
        off = rng.uniform(-0.08, 0.08)
        lp = np.array([x, y0 + gx * x + gz * z, z]) + off * nrm
        world = (quat_rotate(q, lp * sc) + tr).astype(np.float32)
        lp = quat_rotate(qinv, world.astype(np.float64) - tr) / sc
        dist = float((lp - np.array([0.0, y0, 0.0])) @ nrm)
        if abs(dist) < 1e-4:
            continue
        handled, proj, inside = project(v, t, po, world, 0.1)
        assert handled
        lproj = quat_rotate(qinv, proj.astype(np.float64) - tr) / sc
        assert np.allclose(lproj, lp - dist * nrm, atol=3e-5), (lp, dist)
        assert inside == (dist < 0)
        seen[inside] += 1
    assert min(seen.values()) > 100
    # Beyond the query radius nothing is reported.
    far = (quat_rotate(q, np.array([0.0, y0 + 0.5, 0.0]) * sc) + tr).astype(np.float32)
    assert not project(v, t, po, far, 0.1)[0]


def test_bumpy_heightfield_brute_force():
    nr, nc, scale = 11, 13, (1.0, 0.3, 0.9)
    i, j = np.meshgrid(np.arange(nr), np.arange(nc), indexing="ij")
    hgt = np.sin(i / 2.0) * np.cos(j / 3.0)
    v, t = TM.heightfield(hgt, scale)
    po = pose()
    max_dist = 0.08
    rng = np.random.default_rng(11)
    seen = {True: 0, False: 0}
    for _ in range(300):
        x = rng.uniform(-0.5 * scale[0] + max_dist, 0.5 * scale[0] - max_dist)
        z = rng.uniform(-0.5 * scale[2] + max_dist, 0.5 * scale[2] - max_dist)
        y = surface_height(v, t, x, z) + rng.uniform(-0.07, 0.07)
        p = np.array([x, y, z], np.float32)
        d = min(np.linalg.norm(closest_on_triangle(p.astype(np.float64), *(v[tri].astype(np.float64))) - p)
                for tri in t)
        if d > max_dist * 0.9999 or d < 1e-4:
            continue
        handled, proj, inside = project(v, t, po, p, max_dist)
        assert handled
        assert abs(np.linalg.norm(proj.astype(np.float64) - p) - d) < 2e-5
        assert inside == (p[1] < surface_height(v, t, float(p[0]), float(p[2])))
        seen[inside] += 1
    assert min(seen.values()) > 80
