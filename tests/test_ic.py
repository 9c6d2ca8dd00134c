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


def test_hernquist_df_reproduces_density():
    """The distribution function integrates to the model's density
    rho(r) = 1 / (2 pi r (1 + r)^3) (pins the DF formula)."""
    from scipy.integrate import quad

    for r in (0.05, 0.3, 1.0, 4.0, 15.0):
        psi = 1.0 / (1.0 + r)
        vesc = np.sqrt(2.0 * psi)
        rho = quad(lambda v: 4 * np.pi * v * v * ic.hernquist_df(psi - 0.5 * v * v), 0, vesc, epsabs=0, epsrel=1e-10)[0]
        want = 1.0 / (2 * np.pi * r * (1 + r) ** 3)
        assert abs(rho / want - 1) < 1e-6, (r, rho, want)


def test_hernquist_distribution():
    from scipy.integrate import quad

    n = 200000
    p = ic.hernquist(n, 5)
    r = np.linalg.norm(p["pos"], axis=1)
    assert r.max() < 20.0 and abs(p["mass"].sum() - 1.0) < 1e-12
    norm = 400.0 / 441.0
    for x in (0.2, 1.0, 3.0, 10.0):
        assert abs((r < x).mean() - (x * x / (1 + x) ** 2) / norm) < 5e-3
    v = np.linalg.norm(p["vel"], axis=1)
    psi = 1.0 / (1.0 + r)
    assert np.all(v * v < 2 * psi)
    # mean v^2 in radial shells against the DF's second moment
    for a, b in ((0.1, 0.2), (0.9, 1.1), (4.0, 6.0)):
        sel = (r >= a) & (r < b)
        rm = np.median(r[sel])
        ps = 1.0 / (1.0 + rm)
        f = lambda v, k: v ** k * ic.hernquist_df(ps - 0.5 * v * v)
        vmax = np.sqrt(2 * ps)
        want = quad(f, 0, vmax, args=(4,))[0] / quad(f, 0, vmax, args=(2,))[0]
        got = (v[sel] ** 2).mean()
        assert abs(got / want - 1) < 0.03, (a, b, got, want)
    assert np.all(np.abs(p["vel"].mean(0)) < 0.01)


def test_hernquist_deterministic_prefix_stable():
    a = ic.hernquist(500, 9)
    b = ic.hernquist(800, 9)
    assert np.array_equal(a["pos"], b["pos"][:500]) and np.array_equal(a["vel"], b["vel"][:500])
    assert np.array_equal(a, ic.hernquist(500, 9))


def test_hernquist_rejection_bounds_hold():
    """Both rejection regimes are exact: b(q) <= b(1) = 3 pi / 2 on [0, 1]
    (inner proposal), and the tabulated bound dominates g(s) on a fine grid
    for radii in [1, 20) (outer)."""
    q = np.linspace(0.0, 1.0, 100001)
    b = ic._hernquist_b(q)
    assert np.all(b <= 1.5 * np.pi + 1e-12) and abs(b[-1] - 1.5 * np.pi) < 1e-12
    assert np.all(np.diff(b) >= -1e-15)  # f grows with E
    r = np.geomspace(1.0, 19.999, 500)
    psi = 1.0 / (1.0 + r)
    s = np.linspace(0.0, 1.0, 4001)
    g = (s * s * ic.hernquist_df(psi[:, None] * (1.0 - s * s))).max(axis=1)
    bound = ic.hernquist_bound(r)
    assert np.all(g <= bound) and np.all(bound < 2.0 * g)


def test_hernquist_inner_proposal_inverts_cdf():
    """s = t sqrt(r / (1 - t^2)) maps x uniform to density ~ s^2 (r + s^2)^(-5/2)
    on [0, 1]: its CDF s^3 / (r + s^2)^(3/2) * (1 + r)^(3/2) equals x."""
    for r in (1e-6, 0.01, 0.5, 0.999):
        x = np.linspace(0.0, 1.0, 11)
        t = np.cbrt(x) / np.sqrt(1.0 + r)
        s = t * np.sqrt(r / (1.0 - t * t))
        assert np.allclose(s ** 3 / (r + s * s) ** 1.5 * (1 + r) ** 1.5, x, atol=1e-12)
        assert abs(s[-1] - 1.0) < 1e-9  # 1 - t^2 cancels at small r
