"""Every reference citation in the repository (``path/file.jl:N``, ranges
``N-M`` and lists ``N,M-K``) names lines that exist in the reference
checkout, and the positional entry points that the integration table cites
start on the cited line.  Skips where the reference checkout is absent (it
never travels with the repository).

A citation's path is resolved against the reference root, then as a suffix of
a reference path (``flows/update.jl`` -> ``src/flows/update.jl``); a bare file
name must match one file, or, where several files share the name, at least
one of them must be long enough.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
CITE = re.compile(r"(?<![\w/.])((?:[\w.-]+/)*[\w.-]+\.(?:jl|toml|yaml|lcs-yaml|csv|md)):(\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)")
SKIP_DIRS = {".git", "_work", "gpurun_out", "__pycache__", "_build", ".pytest_cache", "oracle/_ref"}
SCAN_EXT = (".py", ".hip", ".h", ".c", ".cpp", ".md", ".jl", ".toml", ".txt")
# top-level documents that this repository does not write (their citations are not checked here)
NOT_OURS = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md"}

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference checkout absent")


def _ref_files():
    out = {}
    for d, dirs, files in os.walk(REF):
        dirs[:] = [x for x in dirs if x != ".git"]
        for f in files:
            p = os.path.relpath(os.path.join(d, f), REF)
            out[p] = None
    return out


def _nlines(p, cache={}):
    if p not in cache:
        with open(os.path.join(REF, p), "rb") as fh:
            cache[p] = fh.read().count(b"\n") + 1
    return cache[p]


def _candidates(path, files):
    if path in files:
        return [path]
    suf = [p for p in files if p.endswith("/" + path)]
    return suf


def _our_files():
    for d, dirs, files in os.walk(ROOT):
        rel = os.path.relpath(d, ROOT)
        dirs[:] = [x for x in dirs if x not in SKIP_DIRS and os.path.join(rel, x).lstrip("./") not in SKIP_DIRS]
        for f in files:
            if rel == "." and f in NOT_OURS:
                continue
            if f.endswith(SCAN_EXT):
                yield os.path.join(d, f)


def _citations():
    for f in _our_files():
        with open(f, encoding="utf-8", errors="replace") as fh:
            for ln, line in enumerate(fh, 1):
                for m in CITE.finditer(line):
                    yield os.path.relpath(f, ROOT), ln, m.group(1), m.group(2)


def test_every_citation_is_in_range():
    files = _ref_files()
    bad = []
    n = 0
    for where, ln, path, spans in _citations():
        cands = _candidates(path, files)
        if not cands:
            continue  # not a reference file (e.g. one of this repo's own)
        n += 1
        top = max(int(x) for x in re.split(r"[-,]", spans))
        if not any(top <= _nlines(c) for c in cands):
            bad.append(f"{where}:{ln}: {path}:{spans} (longest match {max(_nlines(c) for c in cands)} lines)")
    assert n > 100
    assert not bad, "\n".join(bad)


# the positional entry points of the integration table: file, line, the name
# that must appear on that line
ENTRY_POINTS = [
    ("src/buffer.jl", 27, "Buffer"),
    ("src/topologies.jl", 36, "synchalo!"),
    ("src/flows/timestep.jl", 6, "timestep"),
    ("src/particles/timestep.jl", 6, "timestep"),
    ("src/flows/rhs.jl", 9, "rhs!"),
    ("src/flows/update.jl", 12, "update!"),
    ("src/flows/couples/couple.jl", 13, "couple!"),
    ("src/flows/forcings/lf.jl", 6, "forcing!"),
    ("src/flows/forcings/rcf.jl", 1, "forcing!"),
    ("src/particles/boundary.jl", 33, "boundary!"),
    ("src/particles/motion.jl", 15, "motion!"),
    ("src/particles/update.jl", 18, "update!"),
    ("src/particles/makeindex.jl", 9, "makeindex!"),
    ("src/evolve.jl", 15, "evolve!"),
    ("src/evolve.jl", 63, "evolve!"),
    ("src/timestep.jl", 25, "timestep!"),
]


@pytest.mark.parametrize("path,line,name", ENTRY_POINTS)
def test_integration_entry_points_start_on_the_cited_line(path, line, name):
    with open(os.path.join(REF, path), encoding="utf-8") as fh:
        text = fh.read().split("\n")
    assert name in text[line - 1], (path, line, text[line - 1])
    with open(os.path.join(ROOT, "INTEGRATION.md"), encoding="utf-8") as fh:
        integ = fh.read()
    base = path.split("/")[-1]
    starts = set()
    for m in CITE.finditer(integ):
        if m.group(1).split("/")[-1] == base:
            starts |= {int(part.split("-")[0]) for part in m.group(2).split(",")}
    assert line in starts, (path, line, sorted(starts))
