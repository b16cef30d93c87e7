"""Reference citations in the product, the boundary header and the oracle point
at line ranges that exist.

Every `name.ext:a-b` (or `name.ext:a`, or `name.ext:a-b, c-d`) under include/,
oracle/, gauge_higgs_amd/ and tests/ is resolved against the reference tree by
file name (a relative path such as `rep/rep_general.hpp` narrows the match) and
each range must lie inside the file. The reference tree is only present in the
build container; elsewhere the test is skipped."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SCAN = ("include", "oracle", "gauge_higgs_amd", "tests")
EXTS = ("hpp", "cpp", "h", "md", "sh", "dat")
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:%s)):(\d+(?:-\d+)?(?:,\s*\d+(?:-\d+)?)*)" % "|".join(EXTS))
# names of this repository's own files, which are cited by line too
OWN = {"ghx.h", "ghx_api.hip", "kernels.hpp", "device_math.hpp", "learned.hip", "symrep_dev.hpp", "bigrep.hpp",
       "reps.hpp", "cpuref.hpp", "measure.hip", "rep_build.cpp", "series.cpp", "symrep.cpp", "potential.cpp",
       "cli_common.hpp", "ctx.hpp", "launch.hpp", "ops.hpp"}


def reference_files():
    out = {}
    for dp, dn, fn in os.walk(REF):
        dn[:] = [d for d in dn if not d.startswith(".")]
        for f in fn:
            p = os.path.join(dp, f)
            out.setdefault(f, []).append(p)
    return out


def line_count(path, cache={}):
    if path not in cache:
        with open(path, errors="replace") as f:
            cache[path] = sum(1 for _ in f)
    return cache[path]


def citations():
    for top in SCAN:
        for dp, dn, fn in os.walk(os.path.join(ROOT, top)):
            dn[:] = [d for d in dn if d not in ("_build", "__pycache__", "_ref", "bin")]
            for f in fn:
                if not f.endswith((".hpp", ".cpp", ".h", ".hip", ".py", ".md")):
                    continue
                p = os.path.join(dp, f)
                with open(p, errors="replace") as fh:
                    for ln, text in enumerate(fh, 1):
                        for m in CITE.finditer(text):
                            yield os.path.relpath(p, ROOT), ln, m.group(1), m.group(2)


def resolve(name, files):
    base = os.path.basename(name)
    cands = files.get(base, [])
    if "/" in name:
        tail = name.lstrip("./")
        narrowed = [c for c in cands if c.endswith("/" + tail)]
        cands = narrowed or cands
    return cands


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")
def test_reference_citations_point_inside_their_files():
    files = reference_files()
    bad, seen = [], 0
    for where, ln, name, ranges in citations():
        base = os.path.basename(name)
        if base in OWN or name.startswith(("include/", "gauge_higgs_amd/", "oracle/", "tests/")):
            continue
        cands = resolve(name, files)
        if not cands:
            continue  # not a reference file (e.g. a document of this repository)
        seen += 1
        for part in ranges.split(","):
            a, _, b = part.strip().partition("-")
            lo, hi = int(a), int(b or a)
            if not (1 <= lo <= hi and any(hi <= line_count(c) for c in cands)):
                bad.append(f"{where}:{ln}: {name}:{part.strip()} (file has {max(line_count(c) for c in cands)} lines)")
    assert seen > 50
    assert not bad, "\n".join(bad)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")
def test_oracle_functions_cite_the_reference_function_of_the_same_name():
    """A function of the oracle whose header line cites the reference, e.g.
    `void scalar_force() {  // scalar_higgs.hpp:67-99`, restates the reference
    function of that name, so the name appears in the cited range."""
    files = reference_files()
    head = re.compile(r"^\s*(?:[\w:<>&\*]+\s+)+(\w+)\s*\([^)]*\)\s*(?:const\s*)?\{\s*//\s*(.*)$")
    checked, bad = 0, []
    for f in ("cpuref.hpp",):
        path = os.path.join(ROOT, "oracle", f)
        for ln, text in enumerate(open(path), 1):
            m = head.match(text)
            if not m:
                continue
            fn, tail = m.group(1), m.group(2)
            if len(fn) < 5:
                continue  # too short to be told apart from ordinary words
            for c in CITE.finditer(tail):
                cands = resolve(c.group(1), files)
                if not cands:
                    continue
                src = open(cands[0], errors="replace").read().splitlines()
                if not re.search(r"\b%s\b" % fn, "\n".join(src)):
                    continue  # the oracle's own name for the function
                checked += 1
                hit = False
                for part in c.group(2).split(","):
                    a, _, b = part.strip().partition("-")
                    lo, hi = int(a), int(b or a)
                    if re.search(r"\b%s\b" % fn, "\n".join(src[max(0, lo - 3):hi + 1])):
                        hit = True
                if not hit:
                    bad.append(f"oracle/{f}:{ln}: {fn} cites {c.group(0)}")
    assert checked >= 8
    assert not bad, "\n".join(bad)
