"""The drop-in `benchmark` tool (applications/benchmark/benchmark.cpp:72-191):
command-line checks on the CPU, spmv and solve runs on the GPU."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "graphene_amd", "benchmark")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="benchmark tool not built")


def run(*args, timeout=120):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def write(path, obj_text):
    with open(path, "w") as f:
        f.write(obj_text)
    return path


SOLVE_CFG = """// solve benchmark (keys of applications/benchmark/benchmark.cpp:163-183)
{
  "x": 1.0,   /* b = A x */
  "x0": 0.0,
  "benchmark": "solve",
  "solver": {
    "type": "IterativeRefinement", "relResidual": 1e-9, "maxIterations": 10,
    "extendedPrecisionType": "float64",
    "innerSolver": {"type": "PBiCGStab", "maxIterations": 100, "relResidual": 1e-3,
                    "preconditioner": {"type": "ILU"}}
  }
}
"""


def test_cli_usage_errors(tmp_path):
    cfg = write(tmp_path / "c.json", SOLVE_CFG)
    r = run()
    assert r.returncode == 2 and "config is required" in r.stderr
    r = run(cfg, "-p", "4,4,4")
    assert r.returncode == 2 and "--tiles" in r.stderr
    r = run(cfg, "-t", "2")
    assert r.returncode == 2 and "exactly one of --matrix and --poisson" in r.stderr
    r = run(cfg, "-t", "2", "-p", "4,4,4", "-m", cfg)
    assert r.returncode == 2
    r = run(tmp_path / "missing.json", "-t", "2", "-p", "4,4,4")
    assert r.returncode == 2 and "does not exist" in r.stderr
    r = run(cfg, "-t", "2", "-p", "4,4,4", "-d", tmp_path / "nodir")
    assert r.returncode == 2 and "--profile" in r.stderr
    r = run(cfg, "--tiles=2", "--poisson=4,4,4", "--bogus")
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = run(cfg, "--tiles=2", "--poisson=4,4,4", "--partition", tmp_path / "none.txt")
    assert r.returncode == 2 and "--partition" in r.stderr


@pytest.mark.gpu
def test_cli_solve_poisson(tmp_path):
    cfg = write(tmp_path / "c.json", SOLVE_CFG)
    r = run(cfg, "--tiles=4", "--poisson=16,16,16", "-d", tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Done!" in r.stdout
    prof = json.load(open(tmp_path / "profile.json"))
    assert prof["rows"] == 4096 and prof["tiles"] == 4 and prof["converged"]
    assert prof["final_residual"] <= 1e-9 * prof["initial_residual"] * 1.0001
    # x = 1 solves A x = b: the final summary shows ||x|| = sqrt(n) to FP32 accuracy
    line = [l for l in r.stdout.splitlines() if l.startswith("x final")][0]
    norm = float(line.split("= ")[1].split(",")[0])
    assert abs(norm - 64.0) < 1e-4 * 64.0


@pytest.mark.gpu
def test_cli_spmv_matrix_file(tmp_path):
    """-m with a MatrixMarket matrix and x from a MatrixMarket vector file
    (b = A x), then y = A x0: its norm equals the FP64 product's to FP32
    rounding."""
    import scipy.io
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    n = 300
    A = sp.random(n, n, density=0.02, random_state=4, format="csr") + sp.eye(n) * 4.0
    scipy.io.mmwrite(str(tmp_path / "A.mtx"), A)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    scipy.io.mmwrite(str(tmp_path / "x.mtx"), x.reshape(-1, 1))
    cfg = write(tmp_path / "s.json",
                '{"x": "%s", "x0": 0.5, "benchmark": "spmv"}' % (tmp_path / "x.mtx"))
    r = run(cfg, "-t", "3", "-m", tmp_path / "A.mtx", "-d", tmp_path)
    assert r.returncode == 0, r.stderr
    prof = json.load(open(tmp_path / "profile.json"))
    assert prof["rows"] == n and prof["spmv_us"] > 0
    # the spmv benchmark multiplies x = x0 (benchmark.cpp:173-185): y = A (0.5, ...)
    line = [l for l in r.stdout.splitlines() if l.startswith("y:")][0]
    norm = float(line.split("= ")[1].split(",")[0])
    ref = np.linalg.norm(A @ np.full(n, 0.5))
    assert abs(norm - ref) <= 1e-5 * ref


@pytest.mark.gpu
@pytest.mark.parametrize("part", ["graph", "contiguous", "file"])
def test_cli_partitions(tmp_path, part):
    """The three partition sources give the same converged solve (x = 1)."""
    cfg = write(tmp_path / "c.json", SOLVE_CFG)
    arg = part
    if part == "file":
        arg = tmp_path / "p.txt"
        np.savetxt(arg, (np.arange(12 ** 3) // 300).astype(int), fmt="%d")
    r = run(cfg, "-t", "6", "-p", "12,12,12", "--partition", arg, "-d", tmp_path)
    assert r.returncode == 0, r.stderr
    prof = json.load(open(tmp_path / "profile.json"))
    assert prof["tiles"] == 6 and prof["converged"]


@pytest.mark.gpu
def test_cli_unknown_benchmark(tmp_path):
    cfg = write(tmp_path / "u.json", '{"x": 1.0, "benchmark": "nope"}')
    r = run(cfg, "-t", "1", "-p", "4,4,4")
    assert r.returncode == 1 and "Unknown benchmark: nope" in r.stderr
