"""Oracle radial basis and spherical harmonics against the reference's own
radial code, closed forms and the mathematics (radial.cpp:138-169, 344-379;
cutoffs.cpp:44-70; y_grad.cpp:667; ctests/test_radial.cpp,
ctests/test_y_grad.cpp)."""
import json
import os

import numpy as np
import pytest
from numpy.polynomial import legendre

from conftest import GOLDEN


def _golden():
    with open(os.path.join(GOLDEN, "radial_ref.jsonl")) as f:
        return [json.loads(line) for line in f]


def test_radial_matches_reference_build(orc):
    recs = _golden()
    assert len(recs) == 64
    for rec in recs:
        x, y, z = rec["x"], rec["y"], rec["z"]
        r = np.sqrt(x * x + y * y + z * z)
        g, gx, gy, gz = orc.radial(x, y, z, r, rec["rcut"], rec["N"],
                                   cutoff=rec["cutoff"])
        for name, v in (("g", g), ("gx", gx), ("gy", gy), ("gz", gz)):
            np.testing.assert_allclose(v, rec[name], rtol=1e-15, atol=1e-300,
                                       err_msg=name)


@pytest.mark.parametrize("cutoff", ["quadratic", "cosine"])
def test_radial_gradient_fd(orc, cutoff):
    # central differences of g_n(r(x)) (test_radial.cpp:285-344)
    x, y, z, rc, N, h = 2.19, 1.23, -0.24, 7.0, 10, 1e-6
    r = np.sqrt(x * x + y * y + z * z)
    g, gx, gy, gz = orc.radial(x, y, z, r, rc, N, cutoff=cutoff)
    for k, gk in enumerate((gx, gy, gz)):
        d = np.zeros(3)
        d[k] = h
        p = np.array([x, y, z]) + d
        m = np.array([x, y, z]) - d
        gp = orc.radial(*p, np.linalg.norm(p), rc, N, cutoff=cutoff)[0]
        gm = orc.radial(*m, np.linalg.norm(m), rc, N, cutoff=cutoff)[0]
        np.testing.assert_allclose((gp - gm) / (2 * h), gk, rtol=1e-7, atol=1e-7)


def test_radial_outside_cutoff_is_zero(orc):
    g, gx, gy, gz = orc.radial(3.0, 0.0, 0.0, 3.0, 2.5, 6)
    assert not np.any(g) and not np.any(gx)


def test_chebyshev_closed_form(orc):
    # g_n = T_n(r / rc) (r - rc)^2 with T_n(cos t) = cos(n t)
    r, rc, N = 3.3, 5.0, 8
    g = orc.radial(r, 0.0, 0.0, r, rc, N)[0]
    t = np.arccos(r / rc)
    np.testing.assert_allclose(g, np.cos(np.arange(N) * t) * (r - rc) ** 2,
                               rtol=1e-13)


def closed_forms(x, y, z):
    """Real spherical harmonics l <= 3 written out from their definitions."""
    r = np.sqrt(x * x + y * y + z * z)
    pi = np.pi
    c = {}
    c[(0, 0)] = 0.5 / np.sqrt(pi)
    s3 = np.sqrt(3 / (4 * pi))
    c[(1, -1)], c[(1, 0)], c[(1, 1)] = s3 * y / r, s3 * z / r, s3 * x / r
    r2 = r * r
    c[(2, -2)] = 0.5 * np.sqrt(15 / pi) * x * y / r2
    c[(2, -1)] = 0.5 * np.sqrt(15 / pi) * y * z / r2
    c[(2, 0)] = 0.25 * np.sqrt(5 / pi) * (3 * z * z - r2) / r2
    c[(2, 1)] = 0.5 * np.sqrt(15 / pi) * x * z / r2
    c[(2, 2)] = 0.25 * np.sqrt(15 / pi) * (x * x - y * y) / r2
    r3 = r2 * r
    c[(3, -3)] = 0.25 * np.sqrt(35 / (2 * pi)) * y * (3 * x * x - y * y) / r3
    c[(3, -2)] = 0.5 * np.sqrt(105 / pi) * x * y * z / r3
    c[(3, -1)] = 0.25 * np.sqrt(21 / (2 * pi)) * y * (5 * z * z - r2) / r3
    c[(3, 0)] = 0.25 * np.sqrt(7 / pi) * z * (5 * z * z - 3 * r2) / r3
    c[(3, 1)] = 0.25 * np.sqrt(21 / (2 * pi)) * x * (5 * z * z - r2) / r3
    c[(3, 2)] = 0.25 * np.sqrt(105 / pi) * (x * x - y * y) * z / r3
    c[(3, 3)] = 0.25 * np.sqrt(35 / (2 * pi)) * x * (x * x - 3 * y * y) / r3
    return c


def test_ylm_closed_forms_l3(orc):
    rng = np.random.default_rng(7)
    for _ in range(20):
        x, y, z = rng.normal(size=3) * 2
        Y = orc.ylm(x, y, z, 3)[0]
        for (l, m), v in closed_forms(x, y, z).items():
            assert abs(Y[l * l + l + m] - v) < 1e-13, (l, m)


def _sphere_quadrature(L):
    t, wt = legendre.leggauss(L + 2)
    nphi = 2 * L + 3
    phi = 2 * np.pi * np.arange(nphi) / nphi
    pts, w = [], []
    for ti, wi in zip(t, wt):
        st = np.sqrt(1 - ti * ti)
        for p in phi:
            pts.append((st * np.cos(p), st * np.sin(p), ti))
            w.append(wi * 2 * np.pi / nphi)
    return np.array(pts), np.array(w)


def test_ylm_orthonormal(orc):
    L = 8
    pts, w = _sphere_quadrature(L)
    Y = np.array([orc.ylm(*p, L)[0] for p in pts])
    gram = (Y * w[:, None]).T @ Y
    np.testing.assert_allclose(gram, np.eye((L + 1) ** 2), atol=1e-12)


def test_ylm_addition_theorem(orc):
    # sum_m Y_lm(u) Y_lm(v) = (2l+1)/(4 pi) P_l(u.v) (rotation invariance,
    # test_y_grad.cpp:187-276, in closed form)
    rng = np.random.default_rng(3)
    L = 10
    for _ in range(10):
        u, v = rng.normal(size=3), rng.normal(size=3)
        Yu, Yv = orc.ylm(*u, L)[0], orc.ylm(*v, L)[0]
        cosg = u @ v / np.linalg.norm(u) / np.linalg.norm(v)
        for l in range(L + 1):
            sl = slice(l * l, (l + 1) ** 2)
            coef = np.zeros(l + 1)
            coef[l] = 1
            want = (2 * l + 1) / (4 * np.pi) * legendre.legval(cosg, coef)
            assert abs(Yu[sl] @ Yv[sl] - want) < 1e-12


def test_ylm_gradient_fd(orc):
    # test_y_grad.cpp:141-185 point, l = 10, central differences
    p = np.array([2.16, 7.12, -3.14])
    L, h = 10, 1e-6
    _, Yx, Yy, Yz = orc.ylm(*p, L)
    for k, G in enumerate((Yx, Yy, Yz)):
        d = np.zeros(3)
        d[k] = h
        fd = (orc.ylm(*(p + d), L)[0] - orc.ylm(*(p - d), L)[0]) / (2 * h)
        np.testing.assert_allclose(fd, G, atol=1e-9)


def test_ylm_gradient_tangential(orc):
    # Y depends on the direction only: r . grad Y = 0
    rng = np.random.default_rng(5)
    for _ in range(5):
        p = rng.normal(size=3) * 3
        _, Yx, Yy, Yz = orc.ylm(*p, 6)
        np.testing.assert_allclose(p[0] * Yx + p[1] * Yy + p[2] * Yz, 0, atol=1e-12)
