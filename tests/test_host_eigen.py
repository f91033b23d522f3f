"""The host-side eigensolver of the duct's z-wall operator (implicit QL on a
symmetric tridiagonal matrix, dopamine-fdm_amd/csrc/symtri_eigen.h), built on
its own with g++ and checked against LAPACK (scipy's eigh_tridiagonal) and
the eigen identity, on the operators the duct builds (uniform and stretched
z faces) and on matrices with repeated and clustered eigenvalues."""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.linalg

from oracle import ref_numpy as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "dopamine-fdm_amd", "csrc", "symtri_eigen.h")
SRC = r'''
#include "symtri_eigen.h"
extern "C" int eig(int n, const double* d, const double* e, double* lam, double* v) {
    std::vector<double> dd(d, d + n), ee(e, e + (n > 1 ? n - 1 : 0)), vv;
    if (!dfp::symtri_eigen(n, dd, ee, vv)) return 1;
    for (int i = 0; i < n; ++i) lam[i] = dd[i];
    for (size_t i = 0; i < vv.size(); ++i) v[i] = vv[i];
    return 0;
}
'''


@pytest.fixture(scope="module")
def eig():
    if not shutil.which("g++"):
        pytest.skip("no g++")
    d = tempfile.mkdtemp()
    src = os.path.join(d, "eig.cpp")
    with open(src, "w") as f:
        f.write(SRC)
    so = os.path.join(d, "libeig.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.dirname(HDR), src, "-o", so],
                   check=True)
    lib = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.eig.argtypes = [ctypes.c_int, dp, dp, dp, dp]

    def run(dvec, evec):
        n = len(dvec)
        d = np.ascontiguousarray(dvec, dtype=np.float64)
        e = np.ascontiguousarray(evec if n > 1 else [0.0], dtype=np.float64)
        lam = np.zeros(n)
        v = np.zeros(n * n)
        assert lib.eig(n, *(a.ctypes.data_as(dp) for a in (d, e, lam, v))) == 0
        return lam, v.reshape((n, n), order="F")

    yield run
    shutil.rmtree(d, ignore_errors=True)


def _s_matrix(g):
    Dzz = R.dyy_matrix(g.z, g.zg)
    n = Dzz.shape[0]
    w = g.z[1:n + 1] - g.z[:n]
    diag = np.diag(Dzz).copy()
    sub = np.array([1.0 / ((g.zg[k + 1] - g.zg[k]) * np.sqrt(w[k - 1] * w[k])) for k in range(1, n)])
    return diag, sub


@pytest.mark.parametrize("nzm,alpha_z", [(1, 0.0), (2, 0.0), (5, 0.0), (8, 1.8), (64, 2.5), (256, 2.0), (257, 0.0)])
def test_duct_operator_eigen(eig, nzm, alpha_z):
    g = R.duct_grid(4, 3, nzm, 1.0, alpha_z)
    d, e = _s_matrix(g)
    lam, V = eig(d, e)
    n = len(d)
    T = np.diag(d) + (np.diag(e, 1) + np.diag(e, -1) if n > 1 else 0.0)
    scale = max(1.0, np.max(np.abs(lam)))
    assert np.all(np.diff(lam) >= 0)
    assert np.max(np.abs(V.T @ V - np.eye(n))) <= 1e-13 * max(1, n)
    assert np.max(np.abs(T @ V - V * lam[None, :])) <= 1e-13 * scale * max(1, n) ** 0.5
    if n > 1:
        ref = scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True)
        assert np.max(np.abs(lam - ref)) <= 1e-13 * scale


def test_repeated_and_clustered_eigenvalues(eig):
    # a block-diagonal matrix (a zero off-diagonal splits it) with equal blocks
    d = np.array([2.0, 1.0, 2.0, 1.0, 3.0])
    e = np.array([0.5, 0.0, 0.5, 1e-12])
    lam, V = eig(d, e)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.max(np.abs(T @ V - V * lam[None, :])) <= 1e-14 * 4
    assert np.max(np.abs(V.T @ V - np.eye(5))) <= 1e-14 * 5
    assert np.allclose(lam, np.sort(np.linalg.eigvalsh(T)), atol=1e-14)
