"""The C ABI library loads and exports every entry point declared in include/hx3d.h
(no compute calls: these run without a GPU)."""
import os
import re

from horses3d_amd import engine

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(REPO, "include", "hx3d.h")).read()
    return sorted(set(re.findall(r"\b(hx3d_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert sorted(engine.EXPORTS) == declared()


def test_library_exports_every_symbol():
    L = engine.load_library()
    for name in declared():
        assert hasattr(L, name), name
