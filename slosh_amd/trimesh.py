"""Triangle-mesh colliders in the reference's flat GPU layout.

Mirrors ``convert_trimesh_to_gpu`` (src/trimesh.rs:22-103) and
``GpuShape::trimesh`` (src/rbd/shapes.rs:157-173): every trimesh collider
appends to two shared buffers,

* collision vertices (one Vect = x, y, z + pad per entry): the BVH node AABBs
  (mins, maxs per node), the mesh vertices, one pseudo-normal per vertex and
  three edge pseudo-normals per triangle (edges ab, bc, ca);
* collision indices (u32): the BVH topology (entry, exit, shape per node) then
  the triangle vertex indices,

and its shape record holds (vertex root, index root, node count, tag) in ``a``
and (triangle count, vertex count, sampling base id, 0) in ``b`` as u32 bits.

The BVH is a flattened tree walked without a stack (trimesh.slang:111-150):
``entry`` is the node visited when the query does not prune this one
(0xffffffff marks a leaf, whose ``shape`` is its triangle), ``exit`` the node
visited otherwise. The reference builds it with the ``bvh`` crate (not vendored);
any tree with that walk semantics gives the same closest point within the query
radius, so the tree here is a median split of triangle centroids along the
widest axis, flattened in depth-first order.

Pseudo-normals (parry's ``TriMesh::pseudo_normals``, not vendored) follow
Baerentzen and Aanaes, "Signed distance computation using the angle weighted
pseudonormal" (IEEE TVCG 11(3), 2005): per vertex the sum of the incident face
normals weighted by the incident angles, per edge the sum of the two incident
face normals (one for a boundary edge), all normalized. A point is inside when
it lies behind the pseudo-normal of the feature it projects on
(trimesh.slang:152-153), which is exact for closed, consistently oriented meshes.
"""
import numpy as np

from . import layouts as L
from . import models as M

LEAF = 0xFFFFFFFF


class CollisionBuffers:
    """The shared collision vertex / index buffers of a body set."""

    def __init__(self):
        self.vertices = []  # list of (n, 3) float32 chunks
        self.indices = []   # list of u32 chunks
        self.nv = 0
        self.ni = 0

    def _add_vertices(self, v):
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        self.vertices.append(v)
        self.nv += v.shape[0]

    def _add_indices(self, i):
        i = np.ascontiguousarray(i, dtype=np.uint32).reshape(-1)
        self.indices.append(i)
        self.ni += i.shape[0]

    def vertex_array(self):
        """(n, 4) float32: x, y, z, 0 per Vect (the 16-byte storage stride of a vec3)."""
        v = np.concatenate(self.vertices) if self.vertices else np.zeros((0, 3), np.float32)
        out = np.zeros((v.shape[0], 4), np.float32)
        out[:, :3] = v
        return out

    def index_array(self):
        return np.concatenate(self.indices) if self.indices else np.zeros(0, np.uint32)


def _flatten_bvh(tri_min, tri_max, centroids):
    """Median-split BVH over triangles (one triangle per leaf), flattened depth
    first into (aabb mins, aabb maxs, entry, exit, shape) arrays: an inner node's
    entry is its first child (the next node), every node's exit the first node
    after its subtree."""
    mins, maxs, shape, right = [], [], [], []

    def build(ids):
        node = len(mins)
        mins.append(tri_min[ids].min(axis=0))
        maxs.append(tri_max[ids].max(axis=0))
        shape.append(int(ids[0]) if len(ids) == 1 else 0)
        right.append(-1)
        if len(ids) > 1:
            c = centroids[ids]
            axis = int(np.argmax(c.max(axis=0) - c.min(axis=0)))
            order = ids[np.argsort(c[:, axis], kind="stable")]
            half = len(order) // 2
            build(order[:half])
            right[node] = build(order[half:])
        return node

    build(np.arange(tri_min.shape[0]))
    n = len(mins)
    entry = np.full(n, LEAF, np.uint32)
    exit_ = np.zeros(n, np.uint32)
    stack = [(0, n)]
    while stack:
        node, e = stack.pop()
        exit_[node] = e
        if right[node] >= 0:
            entry[node] = node + 1
            stack.append((node + 1, right[node]))
            stack.append((right[node], e))
    return (np.array(mins, np.float32), np.array(maxs, np.float32), entry, exit_, np.array(shape, np.uint32))


def pseudo_normals(vertices, triangles):
    """(vertex pseudo-normals (nv, 3), edge pseudo-normals (nt, 3, 3) for edges ab, bc, ca)."""
    v = np.asarray(vertices, np.float64)
    t = np.asarray(triangles, np.int64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    fn = np.cross(b - a, c - a)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    vn = np.zeros_like(v)
    for k in range(3):
        p, q, r = v[t[:, k]], v[t[:, (k + 1) % 3]], v[t[:, (k + 2) % 3]]
        e1, e2 = q - p, r - p
        cosang = np.einsum("ij,ij->i", e1, e2) / (np.linalg.norm(e1, axis=1) * np.linalg.norm(e2, axis=1))
        ang = np.arccos(np.clip(cosang, -1.0, 1.0))
        np.add.at(vn, t[:, k], fn * ang[:, None])
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    edge_sum = {}
    for ti in range(t.shape[0]):
        for k in range(3):
            key = tuple(sorted((int(t[ti, k]), int(t[ti, (k + 1) % 3]))))
            edge_sum[key] = edge_sum.get(key, 0.0) + fn[ti]
    en = np.zeros((t.shape[0], 3, 3))
    for ti in range(t.shape[0]):
        for k in range(3):
            s = edge_sum[tuple(sorted((int(t[ti, k]), int(t[ti, (k + 1) % 3]))))]
            en[ti, k] = s / max(np.linalg.norm(s), 1e-30)
    return vn.astype(np.float32), en.astype(np.float32)


def convert(vertices, triangles, buffers, vertex_base_id=0):
    """convert_trimesh_to_gpu + GpuShape::trimesh: appends the mesh to `buffers`
    and returns its shape record (layouts.SHAPE)."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
    if t.shape[0] == 0 or t.max() >= v.shape[0]:
        raise ValueError("empty mesh or vertex index out of range")
    tv = v[t]  # (nt, 3, 3)
    bmin, bmax, entry, exit_, shape = _flatten_bvh(tv.min(axis=1), tv.max(axis=1), tv.mean(axis=1))
    vtx_root, idx_root = buffers.nv, buffers.ni
    aabbs = np.empty((2 * bmin.shape[0], 3), np.float32)
    aabbs[0::2], aabbs[1::2] = bmin, bmax
    buffers._add_vertices(aabbs)
    buffers._add_indices(np.stack([entry, exit_, shape], axis=1))
    vn, en = pseudo_normals(v, t)
    buffers._add_vertices(v)
    buffers._add_vertices(vn)
    buffers._add_vertices(en.reshape(-1, 3))
    buffers._add_indices(t)
    sh = np.zeros(1, dtype=L.SHAPE)
    sh["a"][0] = np.array([vtx_root, idx_root, bmin.shape[0], M.SHAPE_TRIMESH], np.uint32).view(np.float32)
    sh["b"][0] = np.array([t.shape[0], v.shape[0], vertex_base_id, 0], np.uint32).view(np.float32)
    return sh[0]


def box_mesh(half_extents):
    """A closed, outward-oriented 12-triangle mesh of the cuboid [-he, he]."""
    hx, hy, hz = half_extents
    v = np.array([[sx * hx, sy * hy, sz * hz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32)
    # vertex id = 4 (x > 0) + 2 (y > 0) + (z > 0)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for q in quads:
        tris += [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
    t = np.array(tris, np.uint32)
    # Orient outward.
    c = v[t].mean(axis=1)
    n = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    flip = np.einsum("ij,ij->i", n, c) < 0
    t[flip] = t[flip][:, ::-1]
    return v, t


def heightfield(heights, scale):
    """A 3D heightfield as the triangle mesh the reference collides with.

    GpuShape::from_parry (src/rbd/shapes.rs:252-268) turns a 3D heightfield into a
    trimesh (parry's ``HeightField::to_trimesh``, then an oriented TriMesh). parry
    is not vendored; this follows its documented layout: ``heights`` is an
    (nrows, ncols) grid over the unit square centred at the origin, rows along z
    and columns along x, scaled by ``scale``, so sample (i, j) sits at
    x = (j / (ncols - 1) - 0.5) sx, y = heights[i, j] sy, z = (i / (nrows - 1) - 0.5) sz.
    Each cell is split along the diagonal from sample (i + 1, j) to (i, j + 1), and
    every triangle faces +y, so points below the surface are inside."""
    hgt = np.asarray(heights, np.float64)
    if hgt.ndim != 2 or min(hgt.shape) < 2:
        raise ValueError("heights must be a 2D grid with at least 2 samples per axis")
    nr, nc = hgt.shape
    sx, sy, sz = (float(c) for c in scale)
    i, j = np.meshgrid(np.arange(nr), np.arange(nc), indexing="ij")
    v = np.stack([(j / (nc - 1) - 0.5) * sx, hgt * sy, (i / (nr - 1) - 0.5) * sz], axis=-1).reshape(-1, 3)
    ci, cj = np.meshgrid(np.arange(nr - 1), np.arange(nc - 1), indexing="ij")
    p00 = (ci * nc + cj).ravel()
    p01, p10, p11 = p00 + 1, p00 + nc, p00 + nc + 1
    t = np.concatenate([np.stack([p00, p10, p01], axis=1), np.stack([p10, p11, p01], axis=1)])
    return v.astype(np.float32), t.astype(np.uint32)


def icosphere(radius, subdivisions=1):
    """A closed, outward-oriented triangulated sphere."""
    p = (1.0 + 5 ** 0.5) / 2
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(x, np.float64) / np.linalg.norm(x) for x in v]
    for _ in range(subdivisions):
        mid = {}

        def m(i, j):
            key = (min(i, j), max(i, j))
            if key not in mid:
                x = v[i] + v[j]
                v.append(x / np.linalg.norm(x))
                mid[key] = len(v) - 1
            return mid[key]
        f = [t for (a, b, c) in f for t in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                                            (m(a, b), m(b, c), m(c, a)))]
    return (np.array(v) * radius).astype(np.float32), np.array(f, np.uint32)
