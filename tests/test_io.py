"""CSV reader and VTK writer (io.cc:94-195 semantics), host only."""
import os

import numpy as np
import pytest

import golden_io
from bhgpu import lib as bl

GOLDEN = golden_io.HERE


def test_reads_reference_data_files():
    p = bl.read_csv(os.path.join(GOLDEN, "two_body.csv"))
    assert len(p) == 2 and list(p["id"]) == [0, 1]
    assert np.array_equal(p["pos"][0], [-0.5, 0, 0]) and np.array_equal(p["vel"][1], [0, 0.1, 0])
    q = bl.read_csv(os.path.join(GOLDEN, "shared_ic_20.csv"))
    rows = np.loadtxt(os.path.join(GOLDEN, "shared_ic_20.csv"), delimiter=",")
    assert len(q) == 20
    assert np.array_equal(q["mass"], rows[:, 1]) and np.array_equal(q["pos"], rows[:, 2:5])


def test_header_blank_lines_crlf(tmp_path):
    f = tmp_path / "a.csv"
    f.write_bytes(b"id,mass,x,y,z,vx,vy,vz\r\n0, 1.5 ,1,2,3,4,5,6\r\n\r\n  \n7,2,-1e-3,0,0,0,0,1e10\n")
    p = bl.read_csv(str(f))
    assert list(p["id"]) == [0, 7] and p["mass"][0] == 1.5 and p["pos"][1][0] == -1e-3 and p["vel"][1][2] == 1e10


def test_malformed_line_names_it(tmp_path):
    f = tmp_path / "b.csv"
    f.write_text("0,1,2,3,4,5,6,7\n1,1,2,3,4,5,6\n")
    with pytest.raises(bl.BHError, match="malformed line 2 of .*'1,1,2,3,4,5,6'"):
        bl.read_csv(str(f))
    f.write_text("0,1,2,3,4,5,6,7\n1,1,2,3,4,5,6,7,8\n")  # a ninth field
    with pytest.raises(bl.BHError, match="malformed line 2 "):
        bl.read_csv(str(f))
    f.write_text("id,m,x,y,z,vx,vy,vz\n\n4294967296,1,2,3,4,5,6,7\n")  # id beyond uint32
    with pytest.raises(bl.BHError, match="malformed line 3 "):
        bl.read_csv(str(f))
    f.write_text("0,1,2,3,4,5,6,7x\n")  # trailing garbage in a number
    with pytest.raises(bl.BHError, match="malformed line 1 "):
        bl.read_csv(str(f))
    with pytest.raises(bl.BHError, match="cannot open"):
        bl.read_csv(str(tmp_path / "missing.csv"))


def test_empty_file(tmp_path):
    f = tmp_path / "c.csv"
    f.write_text("")
    assert len(bl.read_csv(str(f))) == 0


def test_large_file_parallel_chunks(tmp_path):
    n = 700000  # > 32 MiB: parsed in several chunks
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(n, 3))
    f = tmp_path / "big.csv"
    with open(f, "w") as fh:
        fh.write("id,mass,x,y,z,vx,vy,vz\n")
        for i in range(n):
            x, y, z = (float(v) for v in pos[i])
            fh.write(f"{i},{1.0 / n!r},{x!r},{y!r},{z!r},0.0,0.0,0.0\n")
    p = bl.read_csv(str(f))
    assert len(p) == n and np.array_equal(p["id"], np.arange(n)) and np.array_equal(p["pos"], pos)
    # a bad record far into the file is named by its line number
    with open(f, "a") as fh:
        fh.write("1,2,3\n")
    with pytest.raises(bl.BHError, match=f"malformed line {n + 2} of .*'1,2,3'"):
        bl.read_csv(str(f))


def test_vtk_format(tmp_path):
    p = bl.read_csv(os.path.join(GOLDEN, "two_body.csv"))
    out = tmp_path / "s.vtk"
    bl.write_vtk(str(out), p)
    lines = out.read_text().splitlines()
    assert lines[:5] == ["# vtk DataFile Version 3.0", "pepc-ttg particle snapshot", "ASCII", "DATASET POLYDATA",
                         "POINTS 2 double"]
    assert lines[5:7] == ["-0.5 0 0", "0.5 0 0"]
    assert lines[7:10] == ["POINT_DATA 2", "SCALARS mass double 1", "LOOKUP_TABLE default"]
    assert lines[10:12] == ["1", "1"] and lines[12] == "VECTORS velocity double"
    assert lines[13:15] == ["0 -0.1 0", "0 0.1 0"]
