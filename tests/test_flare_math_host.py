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
    L.probe_f16_row_scale.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_int)]
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


def test_f16_row_scale_range(probe):
    """Row scales of the FP16 hi + lo splits stay finite: rows below the
    smallest normal float (zero, denormal, FP64 underflow) are zero rows
    with scale 1; others land in [0.5, 1) (in [1, 2) at the top of the float
    range) with a scale and inverse scale a float holds."""
    e = ctypes.c_int()
    for m in (0.0, 1e-45, 1e-40, 1.1e-38):
        assert probe.probe_f16_row_scale(m, ctypes.byref(e)) == 0 and e.value == 0, m
    for m in (1.17549435e-38, 2.5e-30, 1e-3, 0.5, 0.75, 1.0, 3.0, 1e30, 3.4e38):
        assert probe.probe_f16_row_scale(m, ctypes.byref(e)) == 1, m
        f = np.float32(m) * np.float32(2.0) ** e.value
        assert (0.5 <= f < 1.0 or (m >= 2.0 ** 127 and 1.0 <= f < 2.0)), m
        assert np.isfinite(np.float32(2.0) ** -e.value), m
        assert np.isfinite(np.float32(2.0) ** e.value), m
