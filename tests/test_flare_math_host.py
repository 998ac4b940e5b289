"""The per-pair math of the HIP kernels (flare_math.h), compiled for the host,
against the oracle: Chebyshev x cutoff with its radial derivative and the
real harmonics with gradients up to l = 10."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("probe") / "libprobe.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "flare_pp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "math_probe.cpp"),
                    "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    pd = ctypes.POINTER(ctypes.c_double)
    L.probe_radial.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 2 + [pd, pd]
    L.probe_ylm.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] + [pd] * 4
    L.probe_pair_index.argtypes = [ctypes.c_int] * 3
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("kind,name", [(0, "quadratic"), (1, "cosine")])
def test_radial_matches_oracle(probe, orc, kind, name):
    rng = np.random.default_rng(kind)
    for _ in range(20):
        v = rng.normal(size=3) * 2
        r = np.linalg.norm(v)
        rc, N = 5.5, 12
        g, dg = np.zeros(N), np.zeros(N)
        probe.probe_radial(r, rc, 0.0, N, kind, _p(g), _p(dg))
        go, gx, _, _ = orc.radial(*v, r, rc, N, cutoff=name)
        np.testing.assert_allclose(g, go, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(dg * v[0] / r, gx, rtol=1e-12, atol=1e-12)


def test_ylm_matches_oracle(probe, orc):
    rng = np.random.default_rng(8)
    L = 10
    H = (L + 1) ** 2
    for _ in range(30):
        v = rng.normal(size=3) * 3
        out = [np.zeros(H) for _ in range(4)]
        probe.probe_ylm(*v, L, *[_p(a) for a in out])
        ref = orc.ylm(*v, L)
        for a, b in zip(out, ref):
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-12)


def test_pair_index_layout(probe):
    nr = 6
    seen = []
    for c1 in range(nr):
        for c2 in range(c1, nr):
            seen.append(probe.probe_pair_index(c1, c2, nr))
            assert probe.probe_pair_index(c2, c1, nr) == seen[-1]
    assert seen == list(range(nr * (nr + 1) // 2))  # b2.cpp:209-233 counter
