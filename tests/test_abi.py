"""C-ABI: the library loads without a GPU and exports every symbol that
include/graphene_amd.h declares; host-only entry points work on the CPU."""
import os
import re

import numpy as np

import graphene_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "graphene_amd.h")).read()
    return sorted(set(re.findall(r"\b(gph_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    L = g.lib()
    syms = header_symbols()
    assert len(syms) > 30
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(syms) == g.exported_symbols()


def test_errors_are_reported():
    import pytest
    with pytest.raises(g.GrapheneError) as e:
        g.check_config('{"type": "Nope"}')
    assert "Unknown solver: Nope" in str(e.value)
    with pytest.raises(g.GrapheneError):
        g.HostMatrix.from_coo(3, np.array([0, 5]), np.array([0, 0]), np.ones(2))
