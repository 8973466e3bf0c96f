"""Pins the CPU oracle (CPU only, no GPU).

The reference holds no live tests or fixtures for this path, so each part of the
oracle is pinned by something other than its own formula: brute force on small
inputs, closed forms, invariants, a library routine (LAPACK SVD), or the
known-answer inputs of the reference's commented-out prefix-sum test
(src/grid/prefix_sum.rs:222-231).
"""
import numpy as np
import pytest

from oracle import oracle as O
from slosh_amd import layouts as L
from slosh_amd import models as M
from slosh_amd import scenes

ASSOC = [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)]


# --- keys, hash, scan ---------------------------------------------------------

def test_pack_key_layout():
    # 11 bits x, 10 bits y, 11 bits z with offsets 0x3ff / 0x1ff / 0x3ff (grid.slang:83-90)
    assert O.pack_key(0, 0, 0) == 0x3ff | (0x1ff << 11) | (0x3ff << 21)
    assert O.pack_key(-1023, -511, -1023) == 0
    assert O.pack_key(1024, 512, 1024) == 0xffffffff
    k = O.pack_key(5, -7, 300)
    assert k & 0x7ff == 5 + 0x3ff and (k >> 11) & 0x3ff == 0x1ff - 7 and k >> 21 == 300 + 0x3ff


def test_hash_is_murmur3_mix():
    # k * c1, rotl 15, * c2 (32-bit); checked against Python big-int arithmetic.
    for k in [0, 1, 0xdeadbeef, 0xffffffff, 123456789]:
        x = (k * 0xcc9e2d51) & 0xffffffff
        x = ((x << 15) | (x >> 17)) & 0xffffffff
        x = (x * 0x1b873593) & 0xffffffff
        assert O.hash_key(k) == x


@pytest.mark.parametrize("kind", ["ones", "iota", "random"])
def test_prefix_sum_known_answers(kind):
    # LEN 15071 with all-ones / iota / random%10000 (prefix_sum.rs:222-231); the
    # result is the exclusive scan "as if a 0 was prepended" (prefix_sum.rs:97-109).
    n = 15071
    v = {"ones": np.ones(n), "iota": np.arange(n),
         "random": np.random.default_rng(1).integers(0, 10000, n)}[kind].astype(np.uint32)
    out = O.prefix_sum(v)
    assert out[0] == 0
    if kind == "ones":
        assert np.array_equal(out, np.arange(n, dtype=np.uint32))
    elif kind == "iota":
        i = np.arange(n, dtype=np.uint64)
        assert np.array_equal(out.astype(np.uint64), i * (i - 1) // 2)
    else:
        assert np.array_equal(out, np.concatenate([[0], np.cumsum(v, dtype=np.uint32)[:-1]]))


# --- sparse grid: active blocks and per-block counts by brute force ----------

def _brute_force_blocks(pos, h):
    a = np.rint(pos.astype(np.float32) / np.float32(h)).astype(np.float64) - 1
    prim = np.floor(a / 4).astype(np.int64)
    loc = (a - prim * 4).astype(np.int64)
    table = {}
    for p in {tuple(b) for b in prim}:
        for o in ASSOC:
            table.setdefault((p[0] + o[0], p[1] + o[1], p[2] + o[2]), [0] * 10)
    for b, l in zip(prim, loc):
        table[tuple(b)][l[2]] += 1
        for o in ASSOC[1:]:
            if all(l[i] >= 2 for i in range(3) if o[i] == 1):
                nb = (b[0] + o[0], b[1] + o[1], b[2] + o[2])
                table[nb][4 + max(l[2] - 4 * o[2], -2) + 2] += 1
    return table


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_block_set_and_bucket_counts_brute_force(seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / 32
    pos = rng.uniform(-0.3, 0.4, size=(3000, 3)).astype(np.float32)
    pos[::7] = np.round(pos[::7] / h * 2) * h / 2  # ties exercise round-half-to-even
    parts = M.particle_arrays(pos, h / 4, 1000.0, M.elastic_neo_hookean(1e4, 0.3))
    sc = scenes.Scene("bf", parts, 1e-4, h, 4096, [], [])
    o = O.Oracle(sc)
    o.step(1)
    blocks = o.blocks()
    got = {tuple(int(x) for x in b["vid"]): [int(c) for c in b["bucket_counts"]] for b in blocks}
    want = _brute_force_blocks(pos, h)
    assert got == want
    assert sum(b["num_particles"] for b in blocks) == len(pos)
    for b in blocks:
        assert b["num_particles"] == sum(b["bucket_counts"][:4])
        assert b["num_particles_with_extras"] == sum(b["bucket_counts"])


def test_round_half_to_even_in_block_association():
    # x / h = 2.5 exactly: WGSL round() is half-to-even -> 2, so the associated
    # cell is 1 (block 0); round-half-away would give cell 2.
    h = 0.25
    pos = np.array([[0.625, 0.625, 0.625], [0.875, 0.875, 0.875]], np.float32)  # 2.5 h, 3.5 h
    parts = M.particle_arrays(pos, h / 4, 1000.0, M.elastic_neo_hookean(1e4, 0.3))
    o = O.Oracle(scenes.Scene("tie", parts, 1e-4, h, 64, [], []))
    o.step(1)
    b = {tuple(int(x) for x in r["vid"]): list(r["bucket_counts"]) for r in o.blocks()}
    # both particles land in block (0,0,0): cells 1 (2.5 -> 2 - 1) and 3 (3.5 -> 4 - 1)
    assert b[(0, 0, 0)][:4] == [0, 1, 0, 1]


# --- transfers: conservation and affine reproduction --------------------------

def _free_scene(pos, vel, aff, h=1.0 / 32, model=None, g=(0.0, 0.0, 0.0), dt=1e-4):
    model = model if model is not None else M.elastic_from_lame(M.ELASTIC_NEO_HOOKEAN, 0.0, 0.0)
    parts = M.particle_arrays(pos, h / 4, 1000.0, model, vel)
    parts["kin"]["affine"] = L.mat_to_slang(aff)
    sc = scenes.Scene("free", parts, dt, h, 4096, [], [], gravity=g)
    return sc, parts


def test_p2g_mass_and_momentum_conservation():
    rng = np.random.default_rng(3)
    h = 1.0 / 32
    pos = rng.uniform(0.3, 0.6, size=(4000, 3)).astype(np.float32)
    vel = rng.normal(size=(4000, 3)).astype(np.float32) * 0.01
    aff = rng.normal(size=(4000, 3, 3)).astype(np.float32) * 1e-3
    g = (0.0, -9.81, 0.0)
    sc, parts = _free_scene(pos, vel, aff, h, g=g)
    o = O.Oracle(sc)
    o.step(1)
    nodes = o.nodes().astype(np.float64)
    m = parts["kin"]["mass"].astype(np.float64)
    assert np.isclose(nodes[:, 3].sum(), m.sum(), rtol=1e-5)
    # grid update: v = (p + m g dt) / m, so sum(m v) = sum(m_p v_p) + M g dt
    # (the affine term cancels: sum_i w_ip (x_i - x_p) = 0 for quadratic B-splines)
    p_tot = (nodes[:, :3] * nodes[:, 3:4]).sum(axis=0)
    want = (m[:, None] * vel).sum(axis=0) + m.sum() * np.array(g) * 1e-4
    assert np.allclose(p_tot, want, rtol=1e-4, atol=1e-9)


def test_affine_field_is_reproduced_exactly():
    """v(x) = M x + b with APIC matrix m M^T (shader convention) is a fixed point of
    P2G -> G2P for quadratic B-splines with D^-1 = 4/h^2: after one substep (no
    gravity, no stress) v_p = M x_p + b and the stored affine is m M^T. A
    transposed operand in P2G or G2P breaks this."""
    rng = np.random.default_rng(4)
    h = 1.0 / 32
    Mx = np.array([[0.1, -0.3, 0.2], [0.25, 0.05, -0.15], [-0.2, 0.3, -0.1]])
    b = np.array([0.02, -0.01, 0.03])
    pos = rng.uniform(0.35, 0.55, size=(6000, 3)).astype(np.float32)
    vel = (pos.astype(np.float64) @ Mx.T + b).astype(np.float32)
    mass = np.float32(1000.0) * (np.float32(2 * h / 4) ** 3)
    aff = np.broadcast_to((mass * Mx.T).astype(np.float32), (len(pos), 3, 3))
    sc, parts = _free_scene(pos, vel, aff, h)
    o = O.Oracle(sc)
    o.step(1)
    k = o.kinematics()
    # Every stencil node's momentum is a mass-weighted sum of the same linear field
    # (v_p + A (x_i - x_p) = M x_i + b), so the reproduction holds for every particle.
    inner = np.ones(len(pos), bool)
    v_expect = pos.astype(np.float64) @ Mx.T + b
    assert np.allclose(k["velocity"][inner], v_expect[inner], atol=2e-6)
    a = k["affine"][inner][..., :3].astype(np.float64) / k["mass"][inner][:, None, None]
    assert np.allclose(a, Mx.T, atol=2e-4)


def test_rigid_rotation_gives_skew_velocity_gradient():
    rng = np.random.default_rng(5)
    h = 1.0 / 32
    w = np.array([0.4, -0.7, 0.2])
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    c = np.array([0.45, 0.45, 0.45])
    pos = rng.uniform(0.35, 0.55, size=(6000, 3)).astype(np.float32)
    vel = ((pos - c) @ W.T).astype(np.float32)
    mass = np.float32(1000.0) * (np.float32(2 * h / 4) ** 3)
    aff = np.broadcast_to((mass * W.T).astype(np.float32), (len(pos), 3, 3))
    sc, _ = _free_scene(pos, vel, aff, h)
    o = O.Oracle(sc)
    o.step(1)
    k = o.kinematics()
    inner = np.ones(len(pos), bool)
    a = k["affine"][inner][..., :3].astype(np.float64) / k["mass"][inner][:, None, None]
    assert np.allclose(a, W.T, atol=2e-4)  # [w]x^T = -[w]x: skew-symmetric


# --- constitutive models -------------------------------------------------------

def _rot(v):
    th = np.linalg.norm(v)
    k = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]]) / th
    return np.eye(3) + np.sin(th) * k + (1 - np.cos(th)) * (k @ k)


@pytest.mark.parametrize("model", [M.elastic(1e5, 0.3), M.elastic_neo_hookean(1e5, 0.3)])
def test_stress_vanishes_at_rest_and_under_rotation(model):
    lam, mu = M.lame_lambda_mu(1e5, 0.3)
    for F in [np.eye(3), _rot(np.array([0.3, -1.1, 0.4])), _rot(np.array([2.0, 0.5, -0.7]))]:
        _, Fo, st = O.model_update(model, F.astype(np.float32))
        assert np.allclose(st, 0.0, atol=1e-6 * float(mu)), st


def test_fixed_corotated_closed_form():
    # F = R diag(s): mul(F^T, F - R) = diag(s (s - 1)) independently of R.
    lam, mu = M.lame_lambda_mu(1e5, 0.3)
    s = np.array([1.1, 0.95, 1.02])
    J = s.prod()
    want = np.diag(2 * mu * s * (s - 1)) + lam * (J - 1) * J * np.eye(3)
    for r in [np.zeros(3), np.array([0.3, -0.2, 0.9])]:
        R = np.eye(3) if not r.any() else _rot(r)
        F = (R @ np.diag(s)).astype(np.float32)
        _, _, st = O.model_update(M.elastic(1e5, 0.3), F)
        assert np.allclose(st, want, rtol=2e-5, atol=2e-5 * float(mu))


def test_neo_hookean_closed_form():
    lam, mu = M.lame_lambda_mu(1e5, 0.3)
    F = np.array([[1.05, 0.02, 0.0], [-0.01, 0.97, 0.03], [0.0, 0.01, 1.01]])
    J = np.linalg.det(F)
    want = mu * F.T @ F + (lam * np.log(J) - mu) * np.eye(3)
    _, _, st = O.model_update(M.elastic_neo_hookean(1e5, 0.3), F.astype(np.float32))
    assert np.allclose(st, want, rtol=1e-4, atol=1e-4 * float(mu))
    assert np.allclose(st, st.T, atol=1e-4 * float(mu))


def test_drucker_prager_tension_resets_to_rotation():
    # Expansion (positive strain trace) -> singular values projected to 1:
    # F' = U V^T and the hardening grows by |strain| (drucker_prager.slang:150-152).
    R = _rot(np.array([0.2, 0.1, -0.3]))
    F = (R @ np.diag([1.05, 1.02, 1.01])).astype(np.float32)
    m, Fo, st = O.model_update(M.sand(1e5, 0.3), F)
    assert np.allclose(Fo, R, atol=1e-5)
    strain = np.log([1.05, 1.02, 1.01])
    state = m[1:4].view(np.float32)
    assert np.isclose(state[1], 1.0 + np.linalg.norm(strain), rtol=1e-5)
    assert np.isclose(state[2], np.log(1.05 * 1.02 * 1.01), rtol=1e-4)
    assert np.isclose(state[0], 1.05 * 1.02 * 1.01, rtol=1e-5)


def test_drucker_prager_inside_cone_is_unchanged():
    F = np.diag([0.999, 0.998, 0.9985]).astype(np.float32)  # compression, tiny deviatoric part
    m0 = M.sand(1e5, 0.3)
    m, Fo, st = O.model_update(m0, F)
    assert np.array_equal(Fo, F) and np.array_equal(m, m0)


def test_drucker_prager_projection_lands_on_the_cone():
    lam, mu = M.lame_lambda_mu(1e5, 0.3)
    s0 = np.array([0.85, 1.08, 1.0])  # compression with a large deviatoric part (gamma > 0)
    F = np.diag(s0).astype(np.float32)
    m0 = M.sand(1e5, 0.3)
    m, Fo, st = O.model_update(m0, F)
    assert not np.allclose(Fo, F)
    s1 = np.sort(np.linalg.svd(Fo.astype(np.float64), compute_uv=False))
    e0, e1 = np.log(np.sort(s0)), np.log(s1)
    assert np.isclose(e1.sum(), e0.sum(), atol=1e-5)  # the projection keeps the volumetric strain
    dev = e1 - e1.mean()
    h0, h1, h2, h3 = (m0[4:8].view(np.float32)).astype(np.float64)
    q = 1.0
    ang = h0 + (h1 * q - h3) * np.exp(-h2 * q)
    alpha = np.sqrt(2 / 3) * 2 * np.sin(ang) / (3 - np.sin(ang))
    yield_value = np.linalg.norm(dev) + (3 * lam + 2 * mu) / (2 * mu) * e1.sum() * alpha
    assert abs(yield_value) < 1e-4


def test_plasticity_disabled_when_lambda_is_zero():
    m0 = M.sand(1e5, 0.3)
    m0[8] = np.float32(0.0).view(np.uint32)  # DruckerPrager.lambda = 0
    F = np.diag([0.9, 1.1, 1.0]).astype(np.float32)
    m, Fo, _ = O.model_update(m0, F)
    assert np.array_equal(Fo, F)


# --- SVD -------------------------------------------------------------------------

@pytest.mark.parametrize("f64", [True, False])
def test_svd_against_lapack(f64):
    rng = np.random.default_rng(6)
    mats = [rng.normal(size=(3, 3)) for _ in range(200)]
    mats += [np.eye(3), np.diag([2.0, 2.0, 1.0]), np.diag([1.0, -1.0, 1.0]), _rot(np.array([0.1, 0.2, 0.3]))]
    mats += [np.eye(3) + 1e-3 * rng.normal(size=(3, 3)) for _ in range(50)]
    tol = 1e-10 if f64 else 2e-5
    for A in mats:
        u, s, v = O.svd3(A, f64=f64)
        scale = max(1.0, np.abs(A).max())
        assert np.allclose(u @ np.diag(s) @ v.T, A, atol=tol * scale * 10)
        assert np.allclose(u @ u.T, np.eye(3), atol=tol * 10) and np.allclose(v @ v.T, np.eye(3), atol=tol * 10)
        assert np.linalg.det(u) > 0 and np.linalg.det(v) > 0
        ref = np.linalg.svd(A, compute_uv=False)
        assert np.allclose(np.abs(s), ref, atol=tol * scale * 10)
        assert s[0] >= s[1] - tol and s[1] >= abs(s[2]) - tol
        assert np.sign(s[2]) == np.sign(np.linalg.det(A)) or abs(s[2]) < 1e-12


# --- grid update / boundary conditions: hand-computed cases ----------------------

def _wall_params(dt=1e-3):
    p = np.zeros(1, L.SIM_PARAMS)[0]
    p["gravity"] = (0.0, 0.0, 0.0)
    p["dt"] = dt
    return p


def test_boundary_conditions_hand_cases():
    h = 0.1
    floor = M.body(M.cuboid((10.0, 0.5, 10.0)), (0.0, -0.5, 0.0))  # top face at y = 0
    p = _wall_params()
    v = np.array([0.3, -0.4, 0.1])
    mom = np.array([*(v * 2.0), 2.0], np.float32)
    # Node at y = 0.05 (within 1.5 h above the face): dist = 0.05 - h < 0
    for ty, want in [(M.STICK, np.zeros(3)), (M.SLIP, np.array([0.3, 0.0, 0.1])),
                     (M.SEPARATE, np.array([0.3, 0.0, 0.1]) / np.hypot(0.3, 0.1) * max(0, np.hypot(0.3, 0.1) - 0.5 * 0.4))]:
        out, nd = O.grid_update_node([floor], [M.boundary_condition(ty, 0.5)], h, p, (0.2, 0.05, 0.3), mom)
        assert np.allclose(nd[:3], [0, 1, 0], atol=1e-6) and np.isclose(nd[3], 0.05 - h, atol=1e-6)
        assert np.allclose(out[:3], want, atol=1e-6), (ty, out)
    # Node inside the floor: normal still points outward (up), dist = -0.2 - h
    out, nd = O.grid_update_node([floor], [M.boundary_condition(M.STICK)], h, p, (0.2, -0.2, 0.3), mom)
    assert np.allclose(nd[:3], [0, 1, 0], atol=1e-6) and np.allclose(out[:3], 0.0)
    # Far from the floor (> 1.5 h): untouched
    out, nd = O.grid_update_node([floor], [M.boundary_condition(M.STICK)], h, p, (0.2, 0.2, 0.3), mom)
    assert np.allclose(out[:3], v, atol=1e-6) and nd[3] == np.float32(1e10)
    # Separate, predictive branch: above the margin but arriving within dt
    p2 = _wall_params(dt=1.0)
    out, nd = O.grid_update_node([floor], [M.boundary_condition(M.SEPARATE, 0.0)], h, p2, (0.2, 0.14, 0.3),
                                 np.array([0.0, -0.1, 0.0, 1.0], np.float32))
    dist = 0.14 - h
    assert np.allclose(out[:3], [0.0, -dist / 1.0, 0.0], atol=1e-6)


def test_grid_velocity_clamp_and_gravity():
    h, dt = 0.1, 1e-2
    p = _wall_params(dt)
    p["gravity"] = (0.0, -9.81, 0.0)
    out, _ = O.grid_update_node([], [], h, p, (0.0, 0.0, 0.0), np.array([100.0, 0.02, -0.1, 1.0], np.float32))
    assert np.allclose(out[:3], [h / dt, 0.02 - 9.81 * dt, -0.1], atol=1e-6)
    out, _ = O.grid_update_node([], [], h, p, (0.0, 0.0, 0.0), np.array([1.0, 1.0, 1.0, 0.0], np.float32))
    assert np.array_equal(out[:3], np.zeros(3, np.float32))  # zero mass -> zero velocity


# --- particle update: divergence handling, rigid bodies ------------------------

def test_inverted_particle_is_disabled():
    sc = scenes.elastic_cube(n_side=8)
    sc.particles["def_grad"][5] = L.mat_to_slang(np.diag([1.0, 1.0, -1.0]))
    o = O.Oracle(sc)
    o.step(1)
    k, F = o.kinematics(), o.def_grad()
    assert k["enabled"][5] == 0 and k["mass"][5] == 0.0
    assert np.array_equal(F[5][:, :3], np.eye(3, dtype=np.float32))
    assert np.all(k["velocity"][5] == 0) and np.all(k["affine"][5] == 0)
    assert np.array_equal(o.positions()[5], sc.particles["pos4"][5])  # position not advected
    assert k["enabled"].sum() == sc.n - 1


def test_rigid_body_integration_closed_form():
    sc = scenes.elastic_cube(n_side=4)
    b = M.body(M.cuboid((0.05, 0.05, 0.05)), (0.2, 0.8, 0.2), linvel=(0.5, 0.0, -0.25), angvel=(0.0, 3.0, 0.0),
               mass=2.0, inv_inertia_diag=(10.0, 10.0, 10.0))
    sc.bodies = np.array([b], dtype=L.BODY)
    sc.materials = np.array([M.boundary_condition(M.SLIP)], dtype=L.BOUNDARY_COND)
    o = O.Oracle(sc)
    o.step(1)
    poses, vels = o.bodies()
    dt = 1e-4
    assert np.allclose(poses["translation_scale"][0, :3], [0.2 + 0.5 * dt, 0.8, 0.2 - 0.25 * dt], atol=1e-7)
    q = poses["rotation"][0]
    assert np.allclose(q, [0.0, np.sin(1.5 * dt), 0.0, np.cos(1.5 * dt)], atol=1e-7)
    assert np.allclose(vels["linear"][0], [0.5, -9.81 * dt, -0.25], atol=1e-7)  # gravity after integration
    assert np.allclose(vels["angular"][0], [0.0, 3.0, 0.0])


def test_fixed_body_does_not_move():
    sc = scenes.elastic_cube(n_side=4)
    o = O.Oracle(sc)
    o.step(3)
    poses, vels = o.bodies()
    assert np.array_equal(poses["translation_scale"], sc.bodies["pose"]["translation_scale"])
    assert np.all(vels["linear"] == 0)


# --- timestep bound (timestep_bound.slang:18-92) -------------------------------

def _kin(mass, velocity=(0.0, 0.0, 0.0), enabled=1):
    k = np.zeros(1, L.KINEMATICS)[0]
    k["mass"] = mass
    k["velocity"] = velocity
    k["enabled"] = enabled
    return k


def _props(volume):
    p = np.zeros(1, L.PROPERTIES)[0]
    p["init_volume"] = volume
    return p


def _F(diag=(1.0, 1.0, 1.0)):
    f = np.zeros((3, 4), np.float32)
    f[0, 0], f[1, 1], f[2, 2] = diag
    return f


def _units(seconds):
    return seconds * 1.0e12


@pytest.mark.parametrize("make", [M.elastic, M.elastic_neo_hookean,
                                  lambda e, nu: M.sand(e, nu, 30.0), lambda e, nu: M.sand(e, nu, 30.0, True)])
def test_timestep_bound_is_the_p_wave_cfl_bound_at_rest(make):
    # At rest with F = I the bound is cfl * h / c with c the P-wave speed
    # sqrt(M / rho), M = E (1 - nu) / ((1 + nu)(1 - 2 nu)) (independent of the
    # Lame split the model stores).
    e, nu, rho, h, vol = 2.0e5, 0.3, 1000.0, 1.0 / 128, 1.0e-7
    got = O.particle_timestep_bound(make(e, nu), _kin(rho * vol), _F(), _props(vol), h)
    m = e * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    want = _units(0.5 * h / np.sqrt(m / rho))
    assert abs(got - want) <= 1e-5 * want


def test_timestep_bound_compression_fast_particles_and_minimum():
    e, nu, rho, h, vol = 1.0e5, 0.25, 1000.0, 1.0 / 64, 1.0e-6
    model = M.elastic_neo_hookean(e, nu)
    rest = O.particle_timestep_bound(model, _kin(rho * vol), _F(), _props(vol), h)
    # Compressed by J = 0.81: the current density rho0 / J grows, so with fixed
    # moduli the sound speed drops by sqrt(J) = 0.9 and the bound grows by 1 / 0.9.
    comp = O.particle_timestep_bound(model, _kin(rho * vol), _F((0.9, 0.9, 1.0)), _props(vol), h)
    assert abs(comp / rest - 1.0 / 0.9) < 1e-5
    # Faster than sound: cfl * h / |v| (the velocity bound h / |v| is larger).
    v = 100.0
    fast = O.particle_timestep_bound(model, _kin(rho * vol, (0.0, v, 0.0)), _F(), _props(vol), h)
    assert abs(fast - _units(0.5 * h / v)) <= 1e-5 * _units(0.5 * h / v)
    # Unknown material tag: bound 0 (default.slang:124-125); huge bound saturates.
    assert O.particle_timestep_bound(np.array([7] + [0] * 12, np.uint32), _kin(1.0), _F(), _props(1.0), h) == 0
    assert O.particle_timestep_bound(model, _kin(1.0e12), _F(), _props(1.0), 10.0) == 0xFFFFFFFF
    # Over a scene: the minimum over enabled particles.
    sc = scenes.elastic_cube(n_side=4)
    sc.particles["kin"]["velocity"][5] = (0.0, 50.0, 0.0)
    sc.particles["kin"]["velocity"][9] = (0.0, 400.0, 0.0)
    sc.particles["kin"]["enabled"][9] = 0
    orc = O.Oracle(sc)
    each = [O.particle_timestep_bound(sc.particles["models"][i], sc.particles["kin"][i], sc.particles["def_grad"][i],
                                      sc.particles["props"][i], sc.cell_width)
            for i in range(sc.n) if sc.particles["kin"]["enabled"][i]]
    assert orc.timestep_bound_raw() == min(each)
    assert abs(orc.timestep_bound() - 0.5 * float(sc.cell_width) / 50.0) < 1e-6 * 0.5 * float(sc.cell_width) / 50.0


# --- P2G formulations ----------------------------------------------------------

@pytest.mark.parametrize("make", [lambda: scenes.stir(scenes.elastic_cube(n_side=12), 3),
                                  lambda: scenes.mixed_test_scene(n_side=10),
                                  lambda: scenes.sand_spheres(n_sand=8192, n_per_sphere=1024, n_spheres=4,
                                                              grid_capacity=1 << 14)])
def test_p2g_scatter_is_bit_identical_to_the_shader_gather(make):
    """The oracle's default block-local scatter P2G against the literal per-node
    gather of p2g_scatter_style.slang:41-155 (chunked, z-slab culled): whole
    trajectories must agree bit for bit."""
    sc = make()
    states = []
    for gather in (True, False):
        O.set_p2g_gather(gather)
        try:
            o = O.Oracle(sc)
            o.step(4)
            states.append((o.positions(), o.kinematics(), o.def_grad(), o.nodes()))
        finally:
            O.set_p2g_gather(False)
    for a, b in zip(*states):
        assert a.tobytes() == b.tobytes()
