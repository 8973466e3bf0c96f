"""Pins the oracle's particle update (particle_update.slang:75-140) by closed forms.

Each test builds a scene whose substep has a closed-form outcome fixed
by the mathematics, not by the oracle's own formulas:

* isolated particles at rest (no gravity): the grid carries no momentum, so G2P
  gives v = 0 and a zero velocity gradient; the stored APIC matrix must then be
  the stress term alone, -V0 (4 / h^2) dt tau(F), with tau the textbook
  neo-Hookean / fixed-corotated Kirchhoff stress of F;
* an affine velocity field v(x) = M x + b (reproduced exactly by P2G -> G2P for
  quadratic B-splines) with F0 != I: the deformation gradient must follow
  dF/dt = L F with L = M, i.e. F1 = (I + dt M) F0 for the physical (glam) F;
* isolated particles with a uniform velocity: the norm clamp to h / dt, then the
  implicit damping v / (1 + c dt), then advection x1 = x0 + v1 dt;
* fixed particles in the affine field: x unchanged and v = 0, F still updated by
  the field, and the APIC matrix reduced to the stress term.

Each runs on the oracle (CPU) and, marked gpu, on the HIP path through the C ABI.

Matrices cross the record layout as in the reference: row r of the stored
(shader) matrix is glam column r, so a record holds the transpose of the
physical matrix.
"""
import numpy as np
import pytest

from oracle import oracle as O
from slosh_amd import layouts as L
from slosh_amd import models as M
from slosh_amd import scenes

H = 1.0 / 32

# The same closed forms gate the HIP path (through the C ABI) on a GPU.
BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


def _run(backend, sc, request):
    """One substep of `sc` on the oracle or on the HIP path: (kinematics, positions, def_grad)."""
    if backend == "oracle":
        o = O.Oracle(sc)
        o.step(1)
        return o.kinematics(), o.positions(), o.def_grad()
    from slosh_amd import pipeline as P
    pl, st = request.getfixturevalue("gpu_pipeline")
    data = P.MpmData.from_scene(pl, sc)
    pl.launch_step(data, st)
    data.check_errors(st)
    return data.kinematics(st), data.positions(st), data.def_grad(st)


def _isolated_positions(n_side=3, spacing_cells=5, origin=0.3):
    """Particles 5 cells apart: their quadratic stencils (3 nodes per axis) never share a node."""
    g = np.arange(n_side) * spacing_cells * H + origin + 0.37 * H
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


def _scene(pos, models, vel=None, dt=1e-4, h=H):
    parts = M.particle_arrays(pos, h / 4, 1000.0, models, vel)
    return scenes.Scene("update-pin", parts, dt, h, 4096, [], [], gravity=(0.0, 0.0, 0.0)), parts


def _rot(v):
    th = np.linalg.norm(v)
    k = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]]) / th
    return np.eye(3) + np.sin(th) * k + (1 - np.cos(th)) * (k @ k)


def _tau_neo_hookean(F, lam, mu):
    """Kirchhoff stress of the compressible neo-Hookean solid: mu (F F^T - I) + lam ln J I."""
    J = np.linalg.det(F)
    return mu * (F @ F.T - np.eye(3)) + lam * np.log(J) * np.eye(3)


def _tau_fixed_corotated(F, lam, mu):
    """Kirchhoff stress of the fixed-corotated solid: 2 mu (F - R) F^T + lam (J - 1) J I,
    R the rotation of the polar decomposition F = R S."""
    u, _, vt = np.linalg.svd(F)
    R = u @ vt
    J = np.linalg.det(F)
    return 2 * mu * (F - R) @ F.T + lam * (J - 1) * J * np.eye(3)


def _glam(rec):
    """Stored record(s) (..., 3, 4) -> physical matrices (transpose of the shader rows)."""
    return np.swapaxes(np.asarray(rec, np.float64)[..., :3], -1, -2)


def _record(Fg):
    return L.mat_to_slang(np.swapaxes(np.asarray(Fg, np.float64), -1, -2).astype(np.float32))


# --- stress -> APIC matrix ----------------------------------------------------------

@pytest.mark.parametrize("backend", BACKENDS)
def test_stress_enters_apic_matrix_as_minus_volume_inv_d_dt_tau(backend, request):
    """particle_update.slang:136-140: affine = m C - tau V0 (4/h^2) dt. At rest C = 0."""
    rng = np.random.default_rng(11)
    pos = _isolated_positions()
    n = len(pos)
    E, nu = 1.0e5, 0.3
    lam, mu = (float(x) for x in M.lame_lambda_mu(E, nu))
    models = np.empty((n, 13), np.uint32)
    Fg = np.empty((n, 3, 3))
    for i in range(n):
        s = rng.uniform(0.85, 1.15, 3)
        if i % 3 == 0:  # F = diag(s)
            Fg[i] = np.diag(s)
        else:           # F = R diag(s) Q^T: a general (non-symmetric) F with det > 0
            Fg[i] = _rot(rng.normal(size=3)) @ np.diag(s) @ _rot(rng.normal(size=3)).T
        models[i] = M.elastic_neo_hookean(E, nu) if i % 2 == 0 else M.elastic(E, nu)
    dt = 1e-4
    sc, parts = _scene(pos, models, dt=dt)
    parts["def_grad"][:] = _record(Fg)
    k, P1, Fr = _run(backend, sc, request)
    F1 = _glam(Fr)
    assert np.all(k["enabled"] == 1)
    assert np.array_equal(k["velocity"], np.zeros((n, 3), np.float32))  # no momentum on the grid
    assert np.array_equal(P1[:, :3], pos)                                # x + 0 dt
    # C = 0: F unchanged up to the rounding of F + F (0 dt)
    assert np.allclose(F1, Fg, atol=1e-6)
    V0 = parts["props"]["init_volume"].astype(np.float64)
    sk = V0 * (4.0 / H ** 2) * dt
    got = _glam(k["affine"])
    for i in range(n):
        tau = _tau_neo_hookean(F1[i], lam, mu) if i % 2 == 0 else _tau_fixed_corotated(F1[i], lam, mu)
        want = -sk[i] * tau
        scale = np.abs(want).max()
        assert np.allclose(got[i], want, atol=1e-4 * scale), (i, got[i], want)


# --- deformation-gradient update ----------------------------------------------------

def _affine_field_scene(n=6000, seed=4, dt=1e-3, F0g=None, fixed_every=0, stressed_fixed=False):
    rng = np.random.default_rng(seed)
    Mx = np.array([[0.1, -0.3, 0.2], [0.25, 0.05, -0.15], [-0.2, 0.3, -0.1]])
    b = np.array([0.02, -0.01, 0.03])
    pos = rng.uniform(0.35, 0.55, size=(n, 3)).astype(np.float32)
    vel = (pos.astype(np.float64) @ Mx.T + b).astype(np.float32)
    models = np.broadcast_to(M.elastic_from_lame(M.ELASTIC_NEO_HOOKEAN, 0.0, 0.0), (n, 13)).copy()
    fixed = np.zeros(n, bool)
    if fixed_every:
        fixed[::fixed_every] = True
        if stressed_fixed:
            models[fixed] = M.elastic_neo_hookean(1.0e5, 0.3)
    sc, parts = _scene(pos, models, vel, dt=dt)
    mass = parts["kin"]["mass"][0]
    # APIC matrix m C with C the field's velocity gradient (stored as its transpose).
    parts["kin"]["affine"] = L.mat_to_slang(np.broadcast_to((mass * Mx.T).astype(np.float32), (n, 3, 3)))
    if F0g is not None:
        parts["def_grad"][:] = _record(np.broadcast_to(F0g, (n, 3, 3)))
    parts["props"]["fixed"] = fixed.astype(np.uint32)
    return sc, parts, pos, Mx, b, fixed


F0_GENERAL = np.array([[1.10, 0.15, -0.05], [-0.08, 0.92, 0.12], [0.06, -0.10, 1.04]])


@pytest.mark.parametrize("backend", BACKENDS)
def test_deformation_gradient_follows_the_velocity_gradient(backend, request):
    """particle_update.slang:109: def_grad += mul(def_grad, affine * dt) on the stored
    (transposed) matrices is F1 = (I + dt L) F0 for the physical F, with L = M here.
    F0 does not commute with M, so the other product order fails this."""
    dt = 1e-3
    sc, parts, pos, Mx, b, _ = _affine_field_scene(dt=dt, F0g=F0_GENERAL)
    assert np.abs(F0_GENERAL @ Mx - Mx @ F0_GENERAL).max() > 0.05
    k, P1, Fr = _run(backend, sc, request)
    F1 = _glam(Fr)
    rate = (F1 - F0_GENERAL) / dt
    want = Mx @ F0_GENERAL
    # C is reproduced to ~1e-4 (float32 P2G / G2P); F1 - F0 is F32-rounded at ~1e-7 / dt.
    assert np.abs(rate - want).max() < 2e-3, np.abs(rate - want).max()
    # Advection with the reproduced velocity: x1 = x0 + (M x0 + b) dt.
    v = pos.astype(np.float64) @ Mx.T + b
    assert np.allclose(P1[:, :3], pos + v * dt, atol=1e-6)
    # No stress (lambda = mu = 0): the stored matrix is m C.
    a = _glam(k["affine"]) / k["mass"][:, None, None]
    assert np.allclose(a, Mx, atol=2e-4)


# --- velocity clamp and damping ------------------------------------------------------

@pytest.mark.parametrize("backend", BACKENDS)
def test_velocity_norm_clamp_then_implicit_damping_then_advection(backend, request):
    """particle_update.slang:75-90: |v| > h/dt -> v h/dt / |v|; then v / (1 + c dt); x += v dt."""
    pos = _isolated_positions(n_side=2)
    dt = 1e-4
    lim = H / dt
    cases = [  # (velocity in units of h/dt, damping c)
        ((0.9, -0.8, 0.7), 0.0),    # |v| = 1.39 h/dt: clamped
        ((0.3, 0.1, -0.2), 500.0),  # damped by 1 / 1.05
        ((0.9, 0.9, -0.9), 500.0),  # clamped, then damped (|v| ends at h/dt / 1.05)
        ((0.2, 0.0, 0.1), 0.0),     # untouched
        ((-0.95, 0.6, 0.0), 2000.0),
        ((0.0, -0.5, 0.0), 0.0),
        ((0.7, 0.7, 0.7), 0.0),
        ((0.01, 0.02, 0.03), 1.0e4),
    ]
    v0 = np.array([c[0] for c in cases]) * lim
    damping = np.array([c[1] for c in cases])
    model = M.elastic_from_lame(M.ELASTIC_NEO_HOOKEAN, 0.0, 0.0)
    sc, parts = _scene(pos, model, v0.astype(np.float32), dt=dt)
    parts["props"]["damping"] = damping.astype(np.float32)
    k, P1, _ = _run(backend, sc, request)
    nrm = np.linalg.norm(v0, axis=1, keepdims=True)
    clamped = np.where(nrm > lim, v0 / nrm * lim, v0)
    want_v = clamped / (1.0 + damping[:, None] * dt)
    assert np.allclose(k["velocity"], want_v, rtol=1e-5, atol=1e-5 * lim)
    assert np.allclose(P1[:, :3], pos + want_v * dt, atol=2e-7)
    # the clamp acted on the norm, not per component (each component is below h/dt)
    assert np.all(np.abs(v0) < lim) and np.linalg.norm(k["velocity"][0]) < 1.0001 * lim


# --- fixed particles -------------------------------------------------------------------

@pytest.mark.parametrize("backend", BACKENDS)
def test_fixed_particles_keep_position_and_keep_elastic_response(backend, request):
    """particle_update.slang:86-88 and 119-121: a fixed particle has v = 0 before
    advection and C = 0 after the F update, so x stays, F follows the field and the
    stored matrix is the stress term alone."""
    dt = 1e-3
    sc, parts, pos, Mx, b, fixed = _affine_field_scene(dt=dt, F0g=F0_GENERAL, fixed_every=7, stressed_fixed=True)
    k, P1, Fr = _run(backend, sc, request)
    P = P1[:, :3]
    F1 = _glam(Fr)
    assert np.array_equal(P[fixed], pos[fixed])
    assert np.array_equal(k["velocity"][fixed], np.zeros((fixed.sum(), 3), np.float32))
    rate = (F1[fixed] - F0_GENERAL) / dt
    assert np.abs(rate - Mx @ F0_GENERAL).max() < 2e-3
    lam, mu = (float(x) for x in M.lame_lambda_mu(1.0e5, 0.3))
    V0 = float(parts["props"]["init_volume"][0])
    sk = V0 * (4.0 / H ** 2) * dt
    got = _glam(k["affine"][fixed])
    for i, F in enumerate(F1[fixed]):
        want = -sk * _tau_neo_hookean(F, lam, mu)
        assert np.allclose(got[i], want, atol=1e-4 * np.abs(want).max()), (i, got[i], want)
    # the free particles are advected by the field
    v = pos.astype(np.float64) @ Mx.T + b
    assert np.allclose(P[~fixed], pos[~fixed] + v[~fixed] * dt, atol=1e-6)
