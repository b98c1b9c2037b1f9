"""The C++ adapter (include/nbody_solver_hip.hpp) as a drop-in NBodySolver<TsnePoint2D>.

oracle/_ref/adapter_check is compiled in the build container from tests/adapter_check.cpp
against the reference's own headers (NBodySolver<T>, TsnePoint2D, NBodySolverBH and its t-SNE
kernels) and this library. It runs both solvers through the base-class pointer on the same
points and checks the derivative (median 1e-5, p99 1e-4 relative), Z (1e-6) and the level-3
cell boxes of getNodesBufferData (1e-5 absolute).
"""
import json
import os
import subprocess

import pytest

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
