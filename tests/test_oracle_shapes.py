"""Oracle collider projections (Shape::project_point_on_boundary, shape.slang:180-201) for
the capsule (capsule.slang:41-58), cone (cone.slang:61-120) and cylinder
(cylinder.slang:68-128), pinned by closed-form geometry: the inside flag against the
analytic point-in-shape test, the projection on the boundary (analytic signed distance
zero) and |point - projection| equal to the analytic distance to the boundary, at
identity and rotated / scaled poses."""
import numpy as np
import pytest

from oracle import oracle as O
from slosh_amd import layouts as L
from slosh_amd import models as M


def pose(rot=(0.0, 0.0, 0.0, 1.0), t=(0.0, 0.0, 0.0), scale=1.0):
    p = np.zeros(1, dtype=L.SIM3)[0]
    p["rotation"] = rot
    p["translation_scale"] = (*t, scale)
    return p


def quat_rotate_inv(q, v):
    x, y, z, w = q
    qv = np.array([-x, -y, -z])
    t = 2.0 * np.cross(qv, v)
    return v + w * t + np.cross(qv, t)


def seg_closest(a, b, p):
    ab = b - a
    u = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return a + u * ab


def sd_capsule(p, a, b, r):
    return np.linalg.norm(p - seg_closest(a, b, p)) - r


def sd_cylinder(p, hh, r):
    dr = np.hypot(p[0], p[2]) - r
    dy = abs(p[1]) - hh
    if dr > 0 or dy > 0:
        return np.hypot(max(dr, 0.0), max(dy, 0.0))
    return max(dr, dy)


def sd_cone(p, hh, r):
    # 2D cross-section (rho, y): triangle with the lateral edge (0, hh)-(r, -hh) and the
    # base edge (r, -hh)-(0, -hh).
    q = np.array([np.hypot(p[0], p[2]), p[1]])
    apex, rim, base_c = np.array([0.0, hh]), np.array([r, -hh]), np.array([0.0, -hh])
    d = min(np.linalg.norm(q - seg_closest(apex, rim, q)), np.linalg.norm(q - seg_closest(rim, base_c, q)))
    inside = q[1] >= -hh and q[0] <= r * (hh - q[1]) / (2 * hh)
    return -d if inside else d


def check_shape(shape, sd, poses, n=3000, spread=1.0, seed=1):
    rng = np.random.default_rng(seed)
    checked = {"in": 0, "out": 0}
    for po in poses:
        q = np.array(po["rotation"], np.float64)
        t = np.array(po["translation_scale"][:3], np.float64)
        sc = float(po["translation_scale"][3])
        for _ in range(n):
            local = rng.uniform(-spread, spread, 3)
            world = (quat_rotate_inv(np.array([-q[0], -q[1], -q[2], q[3]]), local * sc) + t).astype(np.float32)
            lp = quat_rotate_inv(q, world.astype(np.float64) - t) / sc
            s = sd(lp)
            if abs(s) < 1e-4:  # too close to the boundary to classify
                continue
            handled, proj, inside = O.project_point_on_boundary(shape, po, world)
            assert handled
            assert inside == (s < 0), (lp, s)
            lproj = quat_rotate_inv(q, proj.astype(np.float64) - t) / sc
            assert abs(sd(lproj)) < 2e-5, (lp, lproj, sd(lproj))
            dist = np.linalg.norm(proj.astype(np.float64) - world)
            assert abs(dist - abs(s) * sc) < 3e-5 * max(1.0, sc), (lp, dist, s)
            checked["in" if s < 0 else "out"] += 1
    assert checked["in"] > 100 and checked["out"] > 100


POSES = [pose(), pose(rot=M.quat_from_scaled_axis((0.3, -0.7, 0.4)), t=(0.5, 0.2, -0.1), scale=1.0),
         pose(rot=M.quat_from_scaled_axis((1.2, 0.1, -0.5)), t=(-0.2, 0.4, 0.3), scale=0.5)]


def test_capsule_projection():
    a, b, r = np.array([0.1, -0.3, 0.2]), np.array([-0.2, 0.4, -0.1]), 0.25
    check_shape(M.capsule(a, b, r), lambda p: sd_capsule(p, a, b, r), POSES)


def test_cylinder_projection():
    check_shape(M.cylinder(0.4, 0.3), lambda p: sd_cylinder(p, 0.4, 0.3), POSES, spread=0.8)


def test_cone_projection():
    check_shape(M.cone(0.4, 0.35), lambda p: sd_cone(p, 0.4, 0.35), POSES, spread=0.8)


def test_capsule_point_on_axis_and_unhandled_shapes():
    sh = M.capsule((0.0, -0.5, 0.0), (0.0, 0.5, 0.0), 0.2)
    handled, proj, inside = O.project_point_on_boundary(sh, pose(), np.zeros(3, np.float32))
    assert handled and inside
    assert abs(np.hypot(proj[0], proj[2]) - 0.2) < 1e-6 and abs(proj[1]) < 1e-6
    tri = np.zeros(1, dtype=L.SHAPE)[0]
    tri["a"][3] = np.uint32(M.SHAPE_TRIMESH).view(np.float32)
    assert not O.project_point_on_boundary(tri, pose(), np.zeros(3, np.float32))[0]
