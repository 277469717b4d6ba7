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


CONSUMER = os.path.join(ROOT, "tests", "abi_consumer", "consumer.c")


def build_consumer(out_dir):
    """Compiles the C99 consumer of include/graphene_amd.h against the built
    library (the header alone, no internal includes)."""
    import subprocess
    exe = os.path.join(str(out_dir), "consumer")
    lib_dir = os.path.join(ROOT, "graphene_amd")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
                    "-I", os.path.join(ROOT, "include"), CONSUMER, "-L", lib_dir,
                    "-lgraphene_amd", f"-Wl,-rpath,{lib_dir}", "-lm", "-o", exe], check=True)
    return exe


def test_header_compiles_as_c_and_cxx_and_links(tmp_path):
    import subprocess
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-x", "c++",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "include", "graphene_amd.h")], check=True)
    exe = build_consumer(tmp_path)
    out = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "consumer ok" in out.stdout, out.stderr
