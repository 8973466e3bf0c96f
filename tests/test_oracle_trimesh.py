"""Oracle trimesh collider (collision/trimesh.slang: stackless BVH walk, triangle
projection with feature location, inside by pseudo-normals) on meshes built by
slosh_amd.trimesh (convert_trimesh_to_gpu's layout), pinned by brute force: the
closest distance over all triangles (Ericson's closest point, FP64), the inside
test of a convex polyhedron (behind every face plane), and the query radius."""
import numpy as np

from oracle import oracle as O
from slosh_amd import layouts as L
from slosh_amd import models as M
from slosh_amd import trimesh as TM


def pose(rot=(0.0, 0.0, 0.0, 1.0), t=(0.0, 0.0, 0.0), scale=1.0):
    p = np.zeros(1, dtype=L.SIM3)[0]
    p["rotation"] = rot
    p["translation_scale"] = (*t, scale)
    return p


def closest_on_triangle(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + ab * (d1 / (d1 - d3))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + ac * (d2 / (d2 - d6))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def brute(p, v, t):
    best = None
    for tri in t:
        q = closest_on_triangle(p, *(v[tri].astype(np.float64)))
        d = np.linalg.norm(q - p)
        if best is None or d < best:
            best = d
    n = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]).astype(np.float64)
    inside = bool(np.all(np.einsum("ij,ij->i", n, p - v[t[:, 0]]) <= 0))
    return best, inside


def quat_rotate(q, v):
    x, y, z, w = q
    qv = np.array([x, y, z])
    t = 2.0 * np.cross(qv, v)
    return v + w * t + np.cross(qv, t)


def check_mesh(v, t, po, n, spread, max_dist, seed):
    buf = TM.CollisionBuffers()
    shape = TM.convert(v, t, buf)
    mesh = (buf.vertex_array(), buf.index_array())
    rng = np.random.default_rng(seed)
    q = np.array(po["rotation"], np.float64)
    tr = np.array(po["translation_scale"][:3], np.float64)
    sc = float(po["translation_scale"][3])
    qinv = np.array([-q[0], -q[1], -q[2], q[3]])
    counts = {"in": 0, "out": 0, "far": 0}
    for _ in range(n):
        local = rng.uniform(-spread, spread, 3)
        world = (quat_rotate(q, local * sc) + tr).astype(np.float32)
        lp = quat_rotate(qinv, world.astype(np.float64) - tr) / sc
        d, inside = brute(lp, v, t)
        handled, proj, ins = O.project_point_on_boundary(shape, po, world, mesh, max_dist)
        if d > max_dist * 1.0001:
            assert not handled, (lp, d)
            counts["far"] += 1
            continue
        if d > max_dist * 0.9999 or d < 1e-4:
            continue  # at the query radius or on the surface: not classified here
        assert handled
        lproj = quat_rotate(qinv, proj.astype(np.float64) - tr) / sc
        assert abs(np.linalg.norm(lproj - lp) - d) < 2e-5, (lp, d)
        assert ins == inside, (lp, d, inside)
        counts["in" if inside else "out"] += 1
    assert min(counts.values()) > 30, counts


def test_box_mesh_projection():
    v, t = TM.box_mesh((0.3, 0.2, 0.25))
    check_mesh(v, t, pose(M.quat_from_scaled_axis((0.4, -0.3, 0.9)), (0.2, -0.1, 0.3), 1.0), 1500, 0.5, 0.15, 3)


def test_icosphere_projection_scaled_pose():
    v, t = TM.icosphere(0.3, 1)
    check_mesh(v, t, pose(M.quat_from_scaled_axis((-0.5, 0.2, 0.1)), (0.0, 0.4, -0.2), 1.5), 1200, 0.45, 0.12, 4)


def test_bvh_layout_and_pseudo_normals():
    v, t = TM.icosphere(1.0, 2)
    buf = TM.CollisionBuffers()
    sh = TM.convert(v, t, buf)
    a = np.array(sh["a"]).view(np.uint32)
    b = np.array(sh["b"]).view(np.uint32)
    assert a[3] == M.SHAPE_TRIMESH and b[0] == len(t) and b[1] == len(v)
    nodes = a[2]
    idx = buf.index_array()
    assert idx.size == 3 * nodes + 3 * len(t)
    assert buf.vertex_array().shape[0] == 2 * nodes + 2 * len(v) + 3 * len(t)
    # The unpruned walk visits every triangle once.
    topo = idx[:3 * nodes].reshape(-1, 3)
    cur, seen = 0, []
    while cur < nodes:
        if topo[cur, 0] == TM.LEAF:
            seen.append(int(topo[cur, 2]))
            cur = topo[cur, 1]
        else:
            cur = topo[cur, 0]
    assert sorted(seen) == list(range(len(t)))
    vn, en = TM.pseudo_normals(v, t)
    # On a sphere the vertex pseudo-normals are radial and every edge normal points outward.
    assert np.min(np.einsum("ij,ij->i", vn, v / np.linalg.norm(v, axis=1, keepdims=True))) > 0.999
    mid = (v[t] + v[np.roll(t, -1, axis=1)]) / 2
    assert np.min(np.einsum("ijk,ijk->ij", en, mid)) > 0


def test_builders_reject_malformed_input():
    import pytest
    buf = TM.CollisionBuffers()
    v, t = TM.box_mesh((0.1, 0.1, 0.1))
    with pytest.raises(ValueError):
        TM.convert(v, np.zeros((0, 3), np.uint32), buf)  # no triangles
    with pytest.raises(ValueError):
        TM.convert(v, t + 8, buf)  # vertex index out of range
    assert buf.nv == 0 and buf.ni == 0  # nothing appended by a rejected mesh
    with pytest.raises(ValueError):
        TM.heightfield(np.zeros(5), (1, 1, 1))  # not a grid
    with pytest.raises(ValueError):
        TM.heightfield(np.zeros((1, 4)), (1, 1, 1))  # fewer than 2 samples along z
    # Two meshes share the buffers: the second's roots follow the first's entries.
    a = TM.convert(v, t, buf)
    b = TM.convert(*TM.icosphere(0.2, 0), buf)
    ra, rb = np.array(a["a"]).view(np.uint32), np.array(b["a"]).view(np.uint32)
    assert ra[0] == 0 and ra[1] == 0
    assert rb[0] == 2 * ra[2] + 2 * len(v) + 3 * len(t)
    assert rb[1] == 3 * ra[2] + 3 * len(t)
    assert buf.vertex_array().shape[0] == buf.nv and buf.index_array().size == buf.ni
