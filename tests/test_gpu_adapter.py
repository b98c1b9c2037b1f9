"""The C++ adapter (include/nbody_solver_hip.hpp) as a drop-in NBodySolver<TsnePoint2D>.

oracle/_ref/adapter_check is compiled in the build container from tests/adapter_check.cpp
against the reference's own headers (NBodySolver<T>, TsnePoint2D, NBodySolverBH and its t-SNE
kernels) and this library. It runs both solvers through the base-class pointer on the same
points and checks the derivative (median 1e-5, p99 1e-4 relative), Z (1e-6) and the level-3
cell boxes of getNodesBufferData (1e-5 absolute). With the FMM solver selected through the
adapter (the reference's NBodySolverFMM_SYM_MORTON needs boost, absent here) the program writes
its points and sums, and they are compared with the oracle's FP64 FMM on the same points and
caller bounds (median 1e-6, p99 1e-5 relative, Z 1e-10).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import rel_err
from tests.test_gpu_fmm import _oracle_fmm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "adapter_check")


@pytest.mark.parametrize("n,theta", [(20000, 0.5), (5000, 1.0)])
def test_adapter_matches_reference_solver(gpu, n, theta):
    if not os.path.exists(EXE):
        pytest.skip("adapter check not built (needs the reference headers at build time)")
    print(f"[adapter_check] n={n} theta={theta}", flush=True)
    r = subprocess.run([EXE, str(n), str(theta)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["boxes"] > 0 and res["boxes"] == res["box_vertices_ref"]
    assert res["median_rel"] <= 1e-5 and res["p99_rel"] <= 1e-4 and res["z_rel"] <= 1e-6


@pytest.mark.parametrize("n,theta", [(20000, 0.5), (3000, 1.0)])
def test_adapter_fmm_matches_oracle(gpu, tmp_path, n, theta):
    if not os.path.exists(EXE):
        pytest.skip("adapter check not built (needs the reference headers at build time)")
    out = tmp_path / "fmm.bin"
    print(f"[adapter_check fmm] n={n} theta={theta}", flush=True)
    r = subprocess.run([EXE, str(n), str(theta), "fmm", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    m = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    assert m == n and len(raw) == 4 + 8 * n + 16 * n + 8
    pos = np.frombuffer(raw[4:4 + 8 * n], dtype=np.float32)
    der = np.frombuffer(raw[4 + 8 * n:4 + 24 * n], dtype=np.float64)
    z = float(np.frombuffer(raw[-8:], dtype=np.float64)[0])
    bx, by, bz = _oracle_fmm(pos[:n].copy(), pos[n:].copy(), theta)
    e = rel_err(der[:n], der[n:], bx, by)
    assert np.median(e) <= 1e-6 and np.quantile(e, 0.99) <= 1e-5, (np.median(e), e.max())
    assert abs(z - bz) <= 1e-10 * abs(bz)
