"""The C99 consumer of include/graphene_amd.h (tests/abi_consumer/consumer.c)
on the GPU: device matrix, b = A x, PBiCGStab + ILU(0) solve, all through the
public header."""
import subprocess

import pytest

from test_abi import build_consumer

pytestmark = pytest.mark.gpu


def test_c_consumer_on_gpu(tmp_path):
    exe = build_consumer(tmp_path)
    out = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "consumer ok" in out.stdout, out.stderr
