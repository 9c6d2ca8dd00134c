"""The C++ driver (bh_run, main.cc's flags and leapfrog) and the C++ host
interface (include/bh_gpu.hpp) on the GPU, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from bhgpu import ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu
BH_RUN = os.path.join(os.path.dirname(bl.LIB_PATH), "bh_run")


def write_csv(path, p):
    with open(path, "w") as f:
        for q in p:
            vals = [float(q["mass"])] + [float(v) for v in q["pos"]] + [float(v) for v in q["vel"]]
            f.write("%d," % int(q["id"]) + ",".join(repr(v) for v in vals) + "\n")


def test_driver_leapfrog_matches_oracle(tmp_path, orc):
    p = ic.random_box(3000, 5, 5.0, vel_spread=0.05)
    src, out = tmp_path / "in.csv", tmp_path / "out.csv"
    write_csv(src, p)
    steps, dt, eps = 5, 1e-3, 1e-2
    r = subprocess.run([BH_RUN, "--input", str(src), "--steps", str(steps), "--dt", str(dt), "--eps", str(eps),
                        "--max-leaf", "8", "--fp64", "--output-csv", str(out), "--time", "--vtk-prefix",
                        str(tmp_path / "snap")], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "step 4 (synced) E_K=" in r.stdout and "timing summary: evals=6" in r.stdout
    assert (tmp_path / "snap_4.vtk").exists()
    got = bl.read_csv(str(out))
    fp = layout.force_params(0.5, eps * eps, 1.0, True)
    q = p
    for k, d in [(dt / 2, dt)] + [(dt, dt)] * (steps - 1) + [(dt / 2, 0.0)]:
        q, _ = orc.step(q, fp, 8, 256, 32, k, d)
    assert np.array_equal(got["id"], q["id"])
    assert np.abs(got["pos"] - q["pos"]).max() < 1e-12 and np.abs(got["vel"] - q["vel"]).max() < 1e-10
    # energy report matches the oracle's energies of the final state
    last = [l for l in r.stdout.splitlines() if "(synced)" in l][0]
    ek = float(last.split("E_K=")[1].split()[0])
    assert ek == pytest.approx(orc.kinetic_energy(q), rel=1e-5)


def test_driver_random_ics_and_errors(tmp_path):
    r = subprocess.run([BH_RUN, "--n", "500", "--steps", "2", "--energy-every", "1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "generated 500 random particles (seed=1, spread=5, vel_spread=0)" in r.stdout
    r = subprocess.run([BH_RUN, "--group-size", "0", "--n", "10"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "per-node" in r.stderr
    bad = tmp_path / "bad.csv"
    bad.write_text("0,1,0,0,0,0,0\n")
    r = subprocess.run([BH_RUN, "--input", str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "malformed line" in r.stderr


@pytest.mark.parametrize("n", [2, 3, 257, 4000, 4001, 20001])
def test_energies_match_oracle(gpu, orc, n):
    # odd and even sizes: the device sum counts each unordered pair once
    # (partners i + 1 .. i + n/2, the pair n/2 apart only from its lower index)
    p = ic.random_box(n, 3, 5.0, vel_spread=0.3)
    ek, ep = gpu.energies(p)
    assert ek == pytest.approx(orc.kinetic_energy(p), rel=1e-14)
    assert ep == pytest.approx(orc.potential_energy(p), rel=1e-12)
