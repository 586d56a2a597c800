"""The repository's documents (the design notes, the integration guide, the
read-me and the profile summaries) point at real things: every repository
path they write in backquotes exists (globs match something), and every test
they name (test_<name>, a file or a function) is defined under tests/.
oracle/_ref is named only to say why there is no reference build."""
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ABSENT_ON_PURPOSE = {"oracle/_ref"}


def _docs():
    names = ["DESIGN.md", "INTEGRATION.md", "README.md"] + sorted(
        os.path.relpath(p, ROOT) for p in glob.glob(os.path.join(ROOT, "profiles", "*.md")))
    return {n: open(os.path.join(ROOT, n)).read() for n in names}


def test_paths_in_documents_exist():
    missing = []
    for name, text in _docs().items():
        for p in re.findall(r"`((?:oetqf\.jl_amd|oracle|tests|include|profiles)/[A-Za-z0-9_./*-]+)`", text):
            p = p.rstrip(".")
            if p in ABSENT_ON_PURPOSE:
                continue
            full = os.path.join(ROOT, p)
            if not (glob.glob(full) if "*" in p else os.path.exists(full)):
                missing.append((name, p))
    assert missing == []


def test_tests_named_in_documents_exist():
    files = glob.glob(os.path.join(ROOT, "tests", "**", "*.py"), recursive=True)
    stems = {os.path.basename(f)[:-3] for f in files}
    src = "".join(open(f).read() for f in files)
    missing = []
    for name, text in _docs().items():
        for t in set(re.findall(r"\b(test_[A-Za-z0-9_]+)", text)):
            if t not in stems and not re.search(r"def %s" % re.escape(t), src):
                missing.append((name, t))
    assert missing == []
