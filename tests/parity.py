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
