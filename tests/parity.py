"""Comparison helpers shared by the parity tests (HIP path vs the CPU oracle)."""
import numpy as np


def rel_l2(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = max(np.linalg.norm(b), floor)
    if den == 0.0:
        return float(np.linalg.norm(a - b))
    return float(np.linalg.norm(a - b) / den)


def block_table(blocks):
    """{vid tuple: (num_particles, num_particles_with_extras, bucket counts tuple)}"""
    return {tuple(int(v) for v in b["vid"]): (int(b["num_particles"]), int(b["num_particles_with_extras"]),
                                              tuple(int(c) for c in b["bucket_counts"])) for b in blocks}


def assert_blocks_equal(gpu_blocks, ref_blocks):
    g = block_table(gpu_blocks)
    r = block_table(ref_blocks)
    assert len(g) == len(gpu_blocks), "duplicate virtual ids in the HIP block set"
    assert set(g) == set(r), f"active-block sets differ: {len(set(g) ^ set(r))} blocks"
    bad = [k for k in r if g[k] != r[k]]
    assert not bad, f"{len(bad)} blocks with different counts, e.g. {bad[0]}: hip {g[bad[0]]} vs oracle {r[bad[0]]}"


def assert_sort_exact_on_gpu_state(scene, data, pipeline, stream):
    """The grid sort on identical inputs after a trajectory: an oracle restarted
    from the HIP path's current particle state and the HIP path each run one more
    substep; the active-block sets and per-block counts must be bit-exact.

    (After many substeps the two trajectories differ at round-off level, so a
    particle within an ulp of a cell face may sit in different cells; the sort
    itself is compared on the same positions.)"""
    import copy
    from oracle import oracle as O
    pos = data.positions(stream)
    pos[:, 3] = 0.0
    sc = copy.copy(scene)
    sc.particles = {"pos4": pos, "kin": data.kinematics(stream), "def_grad": data.def_grad(stream),
                    "props": data.properties(stream), "models": data.models(stream)}
    orc = O.Oracle(sc)
    pipeline.launch_step(data, stream)
    orc.step(1)
    assert_blocks_equal(data.blocks(stream), orc.blocks())


def state_errors(gpu, ref, x0):
    """Relative L2 errors of x (and displacement), v, affine, F."""
    gp, gk, gf = gpu
    rp, rk, rf = ref
    out = {
        "x": rel_l2(gp[:, :3], rp[:, :3]),
        "dx": rel_l2(gp[:, :3] - x0[:, :3], rp[:, :3] - x0[:, :3]),
        "v": rel_l2(gk["velocity"], rk["velocity"]),
        "affine": rel_l2(gk["affine"][..., :3], rk["affine"][..., :3]),
        "F": rel_l2(gf[..., :3], rf[..., :3]),
        "F-I": rel_l2(gf[..., :3] - np.eye(3, dtype=np.float32), rf[..., :3] - np.eye(3, dtype=np.float32)),
        "enabled_diff": int(np.sum(gk["enabled"] != rk["enabled"])),
    }
    return out


def tol_for(steps):
    """Relative L2 bound after `steps` substeps: 1e-5 after one substep and 1e-3
    after 100 (the parity gates), linear in between."""
    return 1.0e-5 * min(max(int(steps), 1), 100)


def tol_long(steps):
    """The same linear growth beyond 100 substeps (1e-5 per substep), for runs of
    hundreds to thousands of substeps (served requests): two evaluations that
    differ by round-off keep drifting apart, the APIC matrix fastest."""
    return 1.0e-5 * max(int(steps), 1)
