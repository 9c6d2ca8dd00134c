"""Seeded synthetic initial conditions (Plummer, uniform cube, random box).

All draws come from one counter-based generator, specified here so that any
implementation reproduces the same particles bit for bit:

    mix64(z)        = splitmix64 finaliser: z ^= z>>30; z *= 0xBF58476D1CE4E5B9;
                      z ^= z>>27; z *= 0x94D049BB133111EB; z ^= z>>31   (mod 2^64)
    state(seed, s)  = mix64(mix64(seed + 0x9E3779B97F4A7C15) + s * 0x9E3779B97F4A7C15)
    draw(seed,s,c)  = mix64(state(seed, s) + (c + 1) * 0xD1B54A32D192ED03)
    uniform         = (draw >> 11) * 2^-53                    in [0, 1)

Stream s is the particle index; counter c selects the draw within the
particle (fixed slots per purpose, below), so a particle's values do not
depend on how many other particles there are or on rejection in others.

Plummer (G = M = a = 1, reference configs): radius by inverse CDF
r = a / sqrt(u^(-2/3) - 1), redrawn until r < 10 a (slots 0..63); isotropic
position angles (slots 64, 65); speed by Aarseth-Henon-Wielen rejection
(1974): q, y uniform, accept when 0.1*y < q^2 (1-q^2)^3.5 (slot pairs
128+2k, 129+2k), v = q * sqrt(2) * (1 + r^2)^(-1/4); isotropic velocity
angles (slots 1024, 1025). Masses 1/N.

Hernquist (G = M = a = 1, mixed-N ensemble): radius by inverse CDF of
M(<r) = r^2 / (1 + r)^2, r = sqrt(u) / (1 - sqrt(u)), redrawn until r < 20 a
(slots 0..63); isotropic position angles (slots 64, 65); speed from the
isotropic ergodic distribution function (Hernquist 1990, ApJ 356, 359,
eq. 17) by rejection in s = v / v_esc: s, y uniform, accept when
y * B(r) < s^2 f(psi (1 - s^2)) with psi = 1 / (1 + r) and B(r) 1.25 times
the largest value of s^2 f(psi (1 - s^2)) over a 257-point grid of s (slot
pairs 128+2k, 129+2k); isotropic velocity angles (slots 1024, 1025).
Masses 1/N.
"""
import numpy as np

from .layout import PARTICLE

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_STEP = np.uint64(0xD1B54A32D192ED03)


def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z ^ (z >> np.uint64(30))
        z = z * _M1
        z = z ^ (z >> np.uint64(27))
        z = z * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def stream_state(seed, streams):
    """Per-stream generator state for `seed` (streams: uint64 array)."""
    with np.errstate(over="ignore"):
        base = _mix64(np.uint64(seed) + _GOLD)
        return _mix64(base + np.asarray(streams, dtype=np.uint64) * _GOLD)


def uniform(state, counter):
    """Uniform doubles in [0, 1) for draw `counter` of each stream state."""
    with np.errstate(over="ignore"):
        z = _mix64(state + (np.uint64(counter) + np.uint64(1)) * _STEP)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _isotropic(state, c0):
    u1 = uniform(state, c0)
    u2 = uniform(state, c0 + 1)
    cos_t = 2.0 * u1 - 1.0
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    phi = 2.0 * np.pi * u2
    return np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], axis=1)


def plummer(n, seed, rmax=10.0):
    """Plummer sphere, G = M = a = 1, truncated at r < rmax (reference configs 0/2)."""
    idx = np.arange(n, dtype=np.uint64)
    st = stream_state(seed, idx)
    r = np.full(n, np.nan)
    todo = np.arange(n)
    for k in range(64):
        if todo.size == 0:
            break
        u = uniform(st[todo], k)
        with np.errstate(divide="ignore", invalid="ignore"):
            rr = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        ok = (u > 0.0) & (rr < rmax)
        r[todo[ok]] = rr[ok]
        todo = todo[~ok]
    if todo.size:
        raise RuntimeError("plummer: radius rejection did not converge")
    pos = _isotropic(st, 64) * r[:, None]
    q = np.full(n, np.nan)
    todo = np.arange(n)
    for k in range(400):
        if todo.size == 0:
            break
        x = uniform(st[todo], 128 + 2 * k)
        y = uniform(st[todo], 129 + 2 * k)
        ok = 0.1 * y < x * x * (1.0 - x * x) ** 3.5
        q[todo[ok]] = x[ok]
        todo = todo[~ok]
    if todo.size:
        raise RuntimeError("plummer: velocity rejection did not converge")
    speed = q * np.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    vel = _isotropic(st, 1024) * speed[:, None]
    p = np.zeros(n, PARTICLE)
    p["id"] = np.arange(n, dtype=np.uint32)
    p["mass"] = 1.0 / n
    p["pos"] = pos
    p["vel"] = vel
    return p


def uniform_cube(n, seed, lo=0.0, hi=1.0):
    """Uniform positions in [lo, hi)^3, masses 1/N, zero velocities (reference config 1)."""
    st = stream_state(seed, np.arange(n, dtype=np.uint64))
    pos = np.stack([uniform(st, c) for c in range(3)], axis=1) * (hi - lo) + lo
    p = np.zeros(n, PARTICLE)
    p["id"] = np.arange(n, dtype=np.uint32)
    p["mass"] = 1.0 / n
    p["pos"] = pos
    return p


def random_box(n, seed, spread, center=(0.0, 0.0, 0.0), vel_spread=0.0):
    """Uniform positions in center +- spread, masses in [0.5, 1.5), velocities in
    +-vel_spread: the distribution of the reference's test and driver inputs
    (tests/test_bh_vs_bruteforce.cc:16-28, main.cc:97-111), from this module's
    generator instead of std::mt19937."""
    st = stream_state(seed, np.arange(n, dtype=np.uint64))
    p = np.zeros(n, PARTICLE)
    p["id"] = np.arange(n, dtype=np.uint32)
    p["mass"] = 0.5 + uniform(st, 0)
    p["pos"] = np.stack([center[a] + spread * (2.0 * uniform(st, 1 + a) - 1.0) for a in range(3)], axis=1)
    if vel_spread:
        p["vel"] = np.stack([vel_spread * (2.0 * uniform(st, 4 + a) - 1.0) for a in range(3)], axis=1)
    return p


def hernquist_df(E):
    """Isotropic distribution function of the Hernquist model, G = M = a = 1
    (Hernquist 1990, eq. 17), for relative binding energies 0 <= E < 1."""
    E = np.asarray(E, dtype=np.float64)
    q = np.sqrt(np.clip(E, 0.0, 1.0 - 1e-16))
    q2 = q * q
    core = 3.0 * np.arcsin(q) + q * np.sqrt(1.0 - q2) * (1.0 - 2.0 * q2) * (8.0 * q2 * q2 - 8.0 * q2 - 3.0)
    return np.maximum(core, 0.0) / ((1.0 - q2) ** 2.5 * 8.0 * np.sqrt(2.0) * np.pi ** 3)


def hernquist(n, seed, rmax=20.0):
    """Hernquist sphere, G = M = a = 1, truncated at r < rmax, isotropic
    ergodic velocities (mixed-N ensemble)."""
    idx = np.arange(n, dtype=np.uint64)
    st = stream_state(seed, idx)
    r = np.full(n, np.nan)
    todo = np.arange(n)
    for k in range(64):
        if todo.size == 0:
            break
        su = np.sqrt(uniform(st[todo], k))
        with np.errstate(divide="ignore"):
            rr = su / (1.0 - su)
        ok = rr < rmax
        r[todo[ok]] = rr[ok]
        todo = todo[~ok]
    if todo.size:
        raise RuntimeError("hernquist: radius rejection did not converge")
    pos = _isotropic(st, 64) * r[:, None]
    psi = 1.0 / (1.0 + r)
    grid = np.linspace(0.0, 1.0, 257)
    bound = np.empty(n)
    for lo in range(0, n, 1 << 16):
        ps = psi[lo:lo + (1 << 16), None]
        bound[lo:lo + (1 << 16)] = 1.25 * (grid * grid * hernquist_df(ps * (1.0 - grid * grid))).max(axis=1)
    s_v = np.full(n, np.nan)
    todo = np.arange(n)
    for k in range(400):
        if todo.size == 0:
            break
        x = uniform(st[todo], 128 + 2 * k)
        y = uniform(st[todo], 129 + 2 * k)
        ok = y * bound[todo] < x * x * hernquist_df(psi[todo] * (1.0 - x * x))
        s_v[todo[ok]] = x[ok]
        todo = todo[~ok]
    if todo.size:
        raise RuntimeError("hernquist: velocity rejection did not converge")
    speed = s_v * np.sqrt(2.0 * psi)
    vel = _isotropic(st, 1024) * speed[:, None]
    p = np.zeros(n, PARTICLE)
    p["id"] = np.arange(n, dtype=np.uint32)
    p["mass"] = 1.0 / n
    p["pos"] = pos
    p["vel"] = vel
    return p
