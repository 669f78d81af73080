"""Every `file.f90:N` / `file.F90:N-M` citation of the reference in the tracked
sources, tests and documents names a file the reference has and lines that
exist in it.  Skipped where the reference sources are absent."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
CITE = re.compile(r"\b([A-Za-z0-9_]+\.(?:f90|F90)):(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")
# a bare continuation such as "`:26-27`" or ", :59-63" cites the file named before it on the line
BARE = re.compile(r"(?<=[\s(`]):(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")
# files the driver writes, and the survey: not this repository's citations
SKIP = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md", "BASELINE.json"}
EXT = (".py", ".hpp", ".hip", ".h", ".F90", ".f90", ".md", ".sh", ".txt")


def tracked():
    out = subprocess.run(["git", "ls-files"], cwd=ROOT, capture_output=True, text=True, check=True).stdout
    return [f for f in out.split() if f.endswith(EXT) and os.path.basename(f) not in SKIP]


def citations():
    for f in tracked():
        with open(os.path.join(ROOT, f), errors="replace") as fh:
            for lineno, line in enumerate(fh, 1):
                named = [(m.start(), m.group(1)) for m in CITE.finditer(line)]
                for m in CITE.finditer(line):
                    yield f, lineno, m.group(1), m.group(2)
                for m in BARE.finditer(line):
                    before = [name for pos, name in named if pos < m.start()]
                    if before:                       # the nearest file named before it
                        yield f, lineno, before[-1], m.group(1)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources absent")
def test_reference_citations_exist():
    counts = {}
    bad = []
    n = 0
    for f, lineno, name, spans in citations():
        path = os.path.join(REF, name)
        if name not in counts:
            counts[name] = sum(1 for _ in open(path, errors="replace")) if os.path.exists(path) else 0
        total = counts[name]
        for span in spans.split(","):
            a, _, b = span.strip().partition("-")
            lo, hi = int(a), int(b or a)
            n += 1
            if total == 0 or lo < 1 or hi < lo or hi > total:
                bad.append(f"{f}:{lineno}: {name}:{span} (file has {total} lines)")
    assert n > 100
    assert not bad, "\n".join(bad)
