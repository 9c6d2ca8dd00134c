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
isotropic ergodic distribution function f(E) = b(q) / (8 sqrt(2) pi^3
(1 - q^2)^(5/2)), q = sqrt(E), b(q) = 3 asin q + q sqrt(1 - q^2) (1 - 2q^2)
(8q^4 - 8q^2 - 3) (Hernquist 1990, ApJ 356, 359, eq. 17), as s = v / v_esc
with density ~ g(s) = s^2 f(psi (1 - s^2)), psi = 1 / (1 + r), drawn by
rejection with slot pairs (128+2k, 129+2k) = (x, y):
  r < 1:  proposal h(s) ~ s^2 (r + s^2)^(-5/2) (g's shape as E -> 1), by
          inversion s = t sqrt(r / (1 - t^2)), t = x^(1/3) / sqrt(1 + r);
          since psi (r + s^2) = 1 - E and b <= b(1) = 3 pi / 2, g / h is
          b(q) / (3 pi / 2) up to a constant: accept when y < b(q) / (3 pi / 2);
  r >= 1: s = x, accept when y B(r) < g(s), B(r) = 1.25 max g over 257
          equally spaced s at the table radius r_j = 20^(j / 4096) <= r,
          j = floor(4096 ln r / ln 20) (f grows with E, so g at r_j
          dominates g at r).
Isotropic velocity angles (slots 1024, 1025). Masses 1/N.
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


def _hernquist_b(q):
    q2 = q * q
    return 3.0 * np.arcsin(q) + q * np.sqrt(1.0 - q2) * (1.0 - 2.0 * q2) * (8.0 * q2 * q2 - 8.0 * q2 - 3.0)


def hernquist_df(E):
    """Isotropic distribution function of the Hernquist model, G = M = a = 1
    (Hernquist 1990, eq. 17), for relative binding energies 0 <= E < 1."""
    E = np.asarray(E, dtype=np.float64)
    q = np.sqrt(np.clip(E, 0.0, 1.0 - 1e-16))
    return np.maximum(_hernquist_b(q), 0.0) / ((1.0 - q * q) ** 2.5 * 8.0 * np.sqrt(2.0) * np.pi ** 3)


_BOUND = []
_NODES = 4096


def hernquist_bound(r):
    """Rejection bound B(r) of the speed draw for radii 1 <= r <= 20."""
    if not _BOUND:
        rj = 20.0 ** (np.arange(_NODES + 1) / _NODES)
        s = np.linspace(0.0, 1.0, 257)
        psi = 1.0 / (1.0 + rj[:, None])
        _BOUND.append(1.25 * (s * s * hernquist_df(psi * (1.0 - s * s))).max(axis=1))
    j = np.floor(_NODES * np.log(np.asarray(r, dtype=np.float64)) / np.log(20.0)).astype(np.int64)
    return _BOUND[0][np.clip(j, 0, _NODES)]


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
    inner = r < 1.0
    bound = np.where(inner, 0.0, hernquist_bound(np.maximum(r, 1.0)))
    s_v = np.full(n, np.nan)
    todo = np.arange(n)
    for k in range(400):
        if todo.size == 0:
            break
        x = uniform(st[todo], 128 + 2 * k)
        y = uniform(st[todo], 129 + 2 * k)
        rt, pt, it = r[todo], psi[todo], inner[todo]
        t = np.cbrt(x) / np.sqrt(1.0 + rt)
        s_in = t * np.sqrt(rt / (1.0 - t * t))
        sv = np.where(it, s_in, x)
        e = pt * (1.0 - sv * sv)
        ok_in = y < _hernquist_b(np.sqrt(np.clip(e, 0.0, 1.0))) / (1.5 * np.pi)
        ok_out = y * bound[todo] < sv * sv * hernquist_df(e)
        ok = np.where(it, ok_in, ok_out)
        s_v[todo[ok]] = sv[ok]
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
