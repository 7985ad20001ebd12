"""Regression fixtures in tests/golden (written by tests/golden/make_golden.py
from the oracle; the reference tree holds no golden vectors for this path):
the oracle must still produce them, and the HIP path must match them to the
parity bars (integer structure bit-exact, matrix and right-hand side within
1e-12, face traces and cell averages within 1e-10 after two steps)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

NAMES = sorted(make_golden.cases())


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def maxrel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    g = load(name)
    got = make_golden.compute(make_golden.cases()[name])
    assert np.array_equal(got["rowptr"], g["rowptr"]) and np.array_equal(got["col"], g["col"])
    assert maxrel(got["val"], g["val"]) < 1e-14 and maxrel(got["rhs"], g["rhs"]) < 1e-14
    assert rel(got["Z_iF"], g["Z_iF"]) < 1e-11 and rel(got["Z_iK"], g["Z_iK"]) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_path_matches_fixture(name):
    from mhfem_amd import capi

    print(f"\n[case] golden {name}", flush=True)
    g = load(name)
    pr = make_golden.cases()[name]
    s = capi.Solver(pr, tol=1e-13)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    assert np.array_equal(rp, g["rowptr"]) and np.array_equal(col, g["col"])
    assert maxrel(val, g["val"]) < 1e-12 and maxrel(rhs, g["rhs"]) < 1e-12
    s.solveLinearSystem()
    s.postIterate(0.0, pr.dt)
    s.step(pr.dt, pr.dt)
    assert rel(s.get_array("Z_iF"), g["Z_iF"]) < 1e-10
    assert rel(s.get_array("Z_iK"), g["Z_iK"]) < 1e-10
    s.close()
