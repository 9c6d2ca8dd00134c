"""Every reference citation (`file.h:a-b`) in the sources and documents lies
inside the cited reference file, and where the citation names a symbol
right before it (`step_particles (integrate_ttg.h:66-69)`), that symbol
appears in the cited lines. Skipped where the reference checkout is absent
(it never travels with the repository)."""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

SCANNED = ["include", "amt-bh-benchmark_amd/csrc", "amt-bh-benchmark_amd/bhgpu", "oracle", "tests", "DESIGN.md",
           "INTEGRATION.md", "README.md", "bench.py", "__graft_entry__.py"]
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:h|cc|cpp|md|py|txt|cmake|sh|csv)):(\d+)(?:-(\d+))?")
# the identifier a citation is attached to: `name (file:a)`, `name file:a`, `name` (file:a)
SYMBOL = re.compile(r"`?([A-Za-z_][A-Za-z0-9_:]*[A-Za-z0-9_])(?:\(\))?`?\s*\(?\s*$")
GENERIC = {"see", "in", "at", "and", "or", "the", "of", "from", "by", "as", "to", "with", "via", "is", "on", "cf",
           "per", "its", "for", "lines", "line", "reference", "here", "ref", "also", "file", "files", "oracle",
           "targets", "root", "quadrupole", "order", "does", "entries", "cell", "walk", "prime", "branch"}


def reference_files():
    out = {}
    for root, dirs, files in os.walk(REF):
        dirs[:] = [d for d in dirs if d != ".git"]
        for f in files:
            path = os.path.join(root, f)
            out.setdefault(f, []).append(os.path.relpath(path, REF))
    return out


def scanned_files():
    for entry in SCANNED:
        path = os.path.join(REPO, entry)
        if os.path.isfile(path):
            yield path
        elif os.path.isdir(path):
            for root, dirs, files in os.walk(path):
                dirs[:] = [d for d in dirs if d not in ("_ref", "__pycache__", "golden")]
                for f in files:
                    if f.endswith((".h", ".hpp", ".hip", ".cpp", ".cc", ".py", ".md")):
                        yield os.path.join(root, f)


def citations():
    refs = reference_files()
    for path in scanned_files():
        if os.path.basename(path) == "test_citations.py":
            continue
        for lineno, line in enumerate(open(path, encoding="utf-8", errors="replace"), 1):
            for m in CITE.finditer(line):
                name = m.group(1)
                base = os.path.basename(name)
                cands = [c for c in refs.get(base, []) if c.endswith(name) or "/" not in name]
                if not cands:
                    continue  # one of this repository's own files
                a = int(m.group(2))
                b = int(m.group(3) or a)
                sym = SYMBOL.search(line[:m.start()])
                yield (os.path.relpath(path, REPO), lineno, name, a, b, cands, sym.group(1) if sym else None)


def _lines(rel):
    return open(os.path.join(REF, rel), encoding="utf-8", errors="replace").read().splitlines()


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
def test_citations_lie_inside_the_cited_files():
    bad = []
    n = 0
    for where, lineno, name, a, b, cands, _ in citations():
        n += 1
        if not any(1 <= a <= b <= len(_lines(c)) for c in cands):
            bad.append(f"{where}:{lineno}: {name}:{a}-{b} (file has {[len(_lines(c)) for c in cands]} lines)")
    assert n > 100, f"only {n} citations found: is the scan broken?"
    assert not bad, "citations outside their files:\n" + "\n".join(bad)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
def test_cited_symbols_appear_in_the_cited_lines():
    bad = []
    for where, lineno, name, a, b, cands, sym in citations():
        if not sym:
            continue
        key = sym.split("::")[-1]
        if key.lower() in GENERIC or len(key) < 4:
            continue
        ok, known = False, False
        for c in cands:
            text = _lines(c)
            if any(key in t for t in text):
                known = True
                if any(key in t for t in text[max(0, a - 3):b + 2]):
                    ok = True
        if known and not ok:
            bad.append(f"{where}:{lineno}: `{key}` not in {name}:{a}-{b}")
    assert not bad, "citations whose named symbol is elsewhere in the file:\n" + "\n".join(bad)
