"""The C-ABI library loads and exports every function include/bh_gpu.h
declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from bhgpu import lib as bl

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "bh_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(bh_[a-z_0-9]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    assert "bh_step" in names and "bh_build_tree" in names and "bh_accel" in names
    assert set(names) == set(bl.SIGNATURES), "bhgpu.lib.SIGNATURES must mirror the header"


def test_library_exports_every_symbol():
    if not os.path.exists(bl.LIB_PATH):
        pytest.fail("libbhgpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(bl.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.bh_max_particles() == (1 << 30) - 1


def test_records_match_header():
    from bhgpu import layout

    assert layout.PARTICLE.itemsize == 64 and layout.PARTICLE.fields["pos"][1] == 16
    assert layout.WALK_NODE.itemsize == 56 and layout.WALK_NODE.fields["first_child"][1] == 40
    assert layout.FORCE_PARAMS.itemsize == 32


def test_struct_layouts_match_c(tmp_path):
    """The Python records have the C compiler's sizes and offsets."""
    import shutil
    import subprocess

    from bhgpu import layout

    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    inc = os.path.dirname(HEADER)
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "bh_gpu.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(bh_particle), sizeof(bh_force_params),\n'
        "         sizeof(bh_system), offsetof(bh_system, kick_dt), sizeof(bh_request), offsetof(bh_request, dt),\n"
        "         sizeof(bh_completion), sizeof(bh_server_stats), sizeof(bh_walk_node), sizeof(bh_step_stats),\n"
        "         offsetof(bh_step_stats, step_ms), offsetof(bh_step_stats, reruns));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c11", "-I", inc, "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [layout.PARTICLE.itemsize, layout.FORCE_PARAMS.itemsize, layout.SYSTEM.itemsize,
            layout.SYSTEM.fields["kick_dt"][1], ctypes.sizeof(bl.Request), bl.Request.dt.offset,
            ctypes.sizeof(bl.Completion), ctypes.sizeof(bl.ServerStats), layout.WALK_NODE.itemsize,
            ctypes.sizeof(bl.StepStats), bl.StepStats.step_ms.offset, bl.StepStats.reruns.offset]
    assert got == want


def test_cpp_wrapper_maps_statuses(tmp_path):
    """bh_gpu.hpp throws as the reference does: an incomplete walk as
    std::logic_error (integrate_ttg.h:104-106), bad input as
    std::invalid_argument, device failures as std::runtime_error."""
    import shutil
    import subprocess

    if not shutil.which("g++"):
        pytest.skip("no g++")
    src = tmp_path / "w.cpp"
    src.write_text(
        '#include <cstdio>\n#include "bh_gpu.hpp"\n'
        "template <class E> int throws(int rc) {\n"
        "  try { bh_gpu::throw_status(rc, \"x\"); } catch (const E&) { return 1; } catch (...) { return 0; }\n"
        "  return 0;\n}\n"
        "int main() {\n"
        "  const int ok = throws<std::logic_error>(BH_ERR_INCOMPLETE) && throws<std::invalid_argument>(BH_ERR_INVALID) &&\n"
        "                 throws<std::invalid_argument>(BH_ERR_NONFINITE) && throws<std::runtime_error>(BH_ERR_HIP) &&\n"
        "                 throws<std::runtime_error>(BH_ERR_OOM);\n"
        "  std::puts(ok ? \"ok\" : \"bad\");\n  return ok ? 0 : 1;\n}\n")
    exe = tmp_path / "w"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.dirname(HEADER), "-I/opt/rocm/include", "-o", str(exe), str(src)],
                   check=True)
    assert subprocess.run([str(exe)], capture_output=True, text=True).stdout.strip() == "ok"
