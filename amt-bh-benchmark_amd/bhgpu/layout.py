"""numpy views of the C-ABI records (include/bh_gpu.h).

The particle and node records keep the reference's in-memory layouts so that
host arrays can be handed to either side unchanged:
  Particle  ttg-parsec/src/particle.h:7-24  (64 B: id, pad, mass, pos[3], vel[3])
  WalkNode  ttg-parsec/src/flat_tree.h:61-70 (56 B)
  ForceParams ttg-parsec/src/traversal_ttg.h:136-141 (theta, eps2, G, use_quadrupole)
"""
import numpy as np

PARTICLE = np.dtype(
    [("id", "<u4"), ("pad", "<u4"), ("mass", "<f8"), ("pos", "<f8", (3,)), ("vel", "<f8", (3,))],
    align=True,
)
assert PARTICLE.itemsize == 64

WALK_NODE = np.dtype(
    [
        ("com", "<f8", (3,)),
        ("mass", "<f8"),
        ("half_width", "<f8"),
        ("first_child", "<u4"),
        ("particle_begin", "<u4"),
        ("particle_count", "<u4"),
        ("num_children", "u1"),
        ("leaf", "u1"),
        ("depth", "u1"),
        ("pad", "u1"),
    ],
    align=True,
)
assert WALK_NODE.itemsize == 56

FORCE_PARAMS = np.dtype(
    [("theta", "<f8"), ("eps2", "<f8"), ("G", "<f8"), ("use_quadrupole", "<i4"), ("pad", "<i4")], align=True
)
assert FORCE_PARAMS.itemsize == 32


SYSTEM = np.dtype(
    [("n", "<u8"), ("fp", FORCE_PARAMS), ("max_leaf", "<u4"), ("group_size", "<u4"), ("target_size", "<u4"),
     ("walk_tile", "<u4"), ("kick_dt", "<f8"), ("drift_dt", "<f8")],
    align=True,
)
assert SYSTEM.itemsize == 72


def system(n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt, walk_tile=0):
    """One bh_system record (walk_tile 0: the context's)."""
    s = np.zeros(1, SYSTEM)
    s["walk_tile"] = walk_tile
    s["n"] = n
    s["fp"] = fp[0] if isinstance(fp, np.ndarray) else fp
    s["max_leaf"], s["group_size"], s["target_size"] = max_leaf, group_size, target_size
    s["kick_dt"], s["drift_dt"] = kick_dt, drift_dt
    return s


def force_params(theta=0.5, eps2=0.0, G=1.0, use_quadrupole=True):
    """ForceParams with the reference's defaults (traversal_ttg.h:136-141)."""
    fp = np.zeros(1, FORCE_PARAMS)
    fp["theta"] = theta
    fp["eps2"] = eps2
    fp["G"] = G
    fp["use_quadrupole"] = 1 if use_quadrupole else 0
    return fp


def make_particles(pos, vel=None, mass=None, ids=None):
    """Packs (n,3) positions (+ velocities, masses, ids) into Particle records."""
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    p = np.zeros(n, PARTICLE)
    p["pos"] = pos
    if vel is not None:
        p["vel"] = vel
    p["mass"] = 1.0 / max(n, 1) if mass is None else mass
    p["id"] = np.arange(n, dtype=np.uint32) if ids is None else ids
    return p
