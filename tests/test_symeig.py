"""The host symmetric eigensolver used to compress the FMM's M2L operators
(aboria_amd/csrc/symeig.hpp), built with g++ and checked against numpy's
eigh and the defining identities: A = V diag(w) V^T, V^T V = I, eigenvalues
descending and equal to numpy's; on random, clustered / repeated-eigenvalue,
rank-deficient (Gram) and diagonal matrices."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "symeig")
        subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp",
                                                                  "symeig_driver.cpp"),
                        "-o", exe], check=True)
        yield exe, d


def _eig(driver, a):
    exe, d = driver
    n = a.shape[0]
    fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
    with open(fi, "wb") as fh:
        fh.write(np.int32(n).tobytes())
        fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    subprocess.run([exe, fi, fo], check=True)
    raw = np.fromfile(fo, dtype=np.float64)
    return raw[:n], raw[n:].reshape(n, n)


def _cases():
    rng = np.random.default_rng(11)
    out = {}
    b = rng.standard_normal((200, 200))
    out["random200"] = b + b.T
    q, _ = np.linalg.qr(rng.standard_normal((120, 120)))
    ev = np.repeat([3.0, 1.0, 1.0, -2.0, 0.5, 0.0], 20)
    out["repeated120"] = (q * ev) @ q.T
    k = rng.standard_normal((150, 4000)) * np.exp(-np.arange(4000) / 300.0)
    out["gram150"] = k @ k.T
    out["diag50"] = np.diag(rng.standard_normal(50))
    out["one"] = np.array([[2.5]])
    out["two"] = np.array([[1.0, 2.0], [2.0, -1.0]])
    return out


@pytest.mark.parametrize("name", list(_cases()))
def test_symeig_against_numpy(driver, name):
    a = _cases()[name]
    w, v = _eig(driver, a)
    n = a.shape[0]
    scale = np.abs(a).max()
    assert np.all(np.diff(w) <= 0)
    assert np.allclose(w, np.sort(np.linalg.eigvalsh(a))[::-1], rtol=0, atol=1e-12 * scale)
    assert np.abs(v.T @ v - np.eye(n)).max() <= 1e-12
    assert np.abs(v @ np.diag(w) @ v.T - a).max() <= 1e-12 * scale
