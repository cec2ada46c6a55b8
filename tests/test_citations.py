"""The reference citations (file:line ranges) in this repository's sources,
headers, tests and documents name lines that exist in the reference, and an
identifier cited with a range ("splatParticles (splat.metal:44-88)") occurs in
that range whenever it occurs in the cited file at all. Runs where the
reference checkout is mounted at /root/reference; skipped elsewhere."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
SCANNED = ["include", "gargantua_amd", "oracle", "tests", "bench.py", "__graft_entry__.py", "DESIGN.md",
           "INTEGRATION.md", "README.md"]
SUFFIXES = (".h", ".hpp", ".cpp", ".hip", ".py", ".md")
CITATION = re.compile(r"([A-Za-z0-9_./-]+\.(?:swift|metal|h|c|md))((?::\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*))")
NAMED = re.compile(r"([A-Za-z_][A-Za-z0-9_]*)\s+\(`?([A-Za-z0-9_./-]+\.(?:swift|metal|h)):(\d+)(?:-(\d+))?")

pytestmark = pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="the reference checkout is not mounted")


def _reference_files():
    out = {}
    for dp, _, fns in os.walk(REFERENCE):
        if ".git" in dp:
            continue
        for f in fns:
            out.setdefault(f, []).append(os.path.relpath(os.path.join(dp, f), REFERENCE))
    return out


def _resolve(files, cited):
    cands = files.get(os.path.basename(cited))
    if not cands:
        return None
    return ([c for c in cands if c.endswith(cited)] or cands)[0]


def _sources():
    for entry in SCANNED:
        p = os.path.join(ROOT, entry)
        if os.path.isfile(p):
            yield p
            continue
        for dp, _, fns in os.walk(p):
            if "_build" in dp or "__pycache__" in dp or os.sep + "bin" in dp:
                continue
            for f in fns:
                if f.endswith(SUFFIXES):
                    yield os.path.join(dp, f)


def _lines(cache, rel):
    if rel not in cache:
        with open(os.path.join(REFERENCE, rel), errors="replace") as f:
            cache[rel] = f.read().splitlines()
    return cache[rel]


def test_every_cited_range_exists():
    files, cache, bad, checked = _reference_files(), {}, [], 0
    for path in _sources():
        text = open(path, errors="replace").read()
        for m in CITATION.finditer(text):
            rel = _resolve(files, m.group(1))
            if rel is None:
                continue
            n = len(_lines(cache, rel))
            for part in m.group(2)[1:].split(","):
                lo, _, hi = part.strip().partition("-")
                checked += 1
                if int(lo) < 1 or int(hi or lo) > n or int(hi or lo) < int(lo):
                    bad.append(f"{os.path.relpath(path, ROOT)}:{text[:m.start()].count(chr(10)) + 1} "
                               f"{m.group(1)}:{part.strip()} ({rel} has {n} lines)")
    assert checked > 250
    assert not bad, "\n".join(bad)


def test_cited_identifiers_lie_in_their_ranges():
    files, cache, bad, checked = _reference_files(), {}, [], 0
    for path in _sources():
        text = open(path, errors="replace").read()
        for m in NAMED.finditer(text):
            rel = _resolve(files, m.group(2))
            if rel is None:
                continue
            lines = _lines(cache, rel)
            name, lo = m.group(1), int(m.group(3))
            hi = int(m.group(4) or lo)
            code_shaped = "_" in name or re.search(r"[a-z][A-Z]", name)
            if not code_shaped or not any(re.search(rf"\b{re.escape(name)}\b", l) for l in lines):
                continue  # a prose word, or a name the cited file does not hold
            checked += 1
            if not any(re.search(rf"\b{re.escape(name)}\b", l) for l in lines[lo - 1:hi]):
                bad.append(f"{os.path.relpath(path, ROOT)}:{text[:m.start()].count(chr(10)) + 1} {name} "
                           f"({m.group(2)}:{lo}-{hi})")
    assert checked > 10
    assert not bad, "\n".join(bad)
