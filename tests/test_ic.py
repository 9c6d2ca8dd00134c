"""Synthetic inputs: the generator is deterministic, follows its stated
counter-based specification, and produces the stated distributions."""
import numpy as np

from bhgpu import ic


def splitmix_ref(seed, stream, counter):
    M = (1 << 64) - 1

    def mix(z):
        z &= M
        z ^= z >> 30
        z = (z * 0xBF58476D1CE4E5B9) & M
        z ^= z >> 27
        z = (z * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    st = mix(mix(seed + 0x9E3779B97F4A7C15) + stream * 0x9E3779B97F4A7C15)
    return (mix(st + (counter + 1) * 0xD1B54A32D192ED03) >> 11) * 2.0 ** -53


def test_uniform_follows_spec():
    st = ic.stream_state(7, np.arange(5, dtype=np.uint64))
    for c in (0, 1, 64, 1025):
        u = ic.uniform(st, c)
        for s in range(5):
            assert u[s] == splitmix_ref(7, s, c)


def test_deterministic_and_prefix_stable():
    a = ic.plummer(1000, 3)
    b = ic.plummer(1000, 3)
    c = ic.plummer(1500, 3)
    assert np.array_equal(a, b)
    assert np.array_equal(a["pos"], c["pos"][:1000])  # a particle does not depend on N
    assert not np.array_equal(a["pos"], ic.plummer(1000, 4)["pos"])


def test_plummer_distribution():
    n = 200000
    p = ic.plummer(n, 1)
    r = np.linalg.norm(p["pos"], axis=1)
    assert r.max() < 10.0 and abs(p["mass"].sum() - 1.0) < 1e-12
    # enclosed-mass profile M(<r) = r^3 / (1 + r^2)^(3/2), renormalised to r < 10
    for x in (0.5, 1.0, 2.0, 5.0):
        want = (x ** 3 / (1 + x * x) ** 1.5) / (1000 / 101 ** 1.5)
        assert abs((r < x).mean() - want) < 5e-3
    v = np.linalg.norm(p["vel"], axis=1)
    vesc = np.sqrt(2.0) * (1 + r * r) ** -0.25
    assert np.all(v < vesc)
    # virial ratio of the untruncated model: 2K/|W| = 1; kinetic energy 3 pi / 64
    ek = 0.5 * (p["mass"] * v * v).sum()
    assert abs(ek - 3 * np.pi / 64) < 0.01
    # isotropy
    assert np.all(np.abs(p["pos"].mean(0)) < 0.02) and np.all(np.abs(p["vel"].mean(0)) < 0.01)


def test_uniform_cube():
    p = ic.uniform_cube(100000, 2)
    assert p["pos"].min() >= 0.0 and p["pos"].max() < 1.0
    assert np.allclose(p["pos"].mean(0), 0.5, atol=5e-3)
    assert np.all(p["vel"] == 0) and abs(p["mass"].sum() - 1) < 1e-12
