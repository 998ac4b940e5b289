"""libflare_hip.so on the CPU: exported symbols and the host-side coefficient
file reader/writer (pair_flare.cpp:263-416; sparse_gp.cpp:1142-1206).
No device work happens here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


@pytest.fixture(scope="module")
def capi():
    from flare_pp_amd import _capi
    _capi.lib()
    return _capi


def header_functions():
    with open(os.path.join(ROOT, "include", "flare_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(flare_\w+)\s*\(",
                                 text, re.M)))


def test_every_header_symbol_is_exported(capi):
    names = header_functions()
    assert len(names) >= 30
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(capi.SIGNATURES), set(names) - set(capi.SIGNATURES)


def test_version(capi):
    assert b"gfx950" in capi.lib().flare_version()


def read(capi, path, values=True):
    info = capi.PairInfo()
    rc = capi.lib().flare_coeff_read(str(path).encode(), ctypes.byref(info), None,
                                     None)
    if rc:
        return rc, None, None, None
    cut = np.zeros(info.n_species ** 2)
    beta = np.zeros(info.n_species * info.beta_size)
    capi.check(capi.lib().flare_coeff_read(
        str(path).encode(), ctypes.byref(info),
        cut.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        beta.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return 0, info, cut, beta


def tokens_after(path, n_header):
    with open(path) as f:
        lines = f.read().splitlines()
    return np.array([float(t) for line in lines[n_header:] for t in line.split()])


def test_legacy_si_fixture(capi):
    # lammps_plugins/test_compute/Si.txt: no power line, 1 species, N 8, L 4
    path = os.path.join(GOLDEN, "Si.txt")
    rc, info, cut, beta = read(capi, path)
    assert rc == 0 and info.legacy == 1 and info.power == 2
    assert (info.n_species, info.n_max, info.l_max, info.beta_size) == (1, 8, 4, 16290)
    assert info.n_descriptors == 180 and info.cutoff == 5.0
    assert info.radial_basis == b"chebyshev" and info.cutoff_name == b"quadratic"
    vals = tokens_after(path, 4)
    assert vals[0] == 5.0 and len(vals) == 1 + 16290
    np.testing.assert_array_equal(beta, vals[1:])
    np.testing.assert_array_equal(cut, [5.0])


def test_legacy_two_species_cosine_fixture(capi):
    path = os.path.join(GOLDEN, "beta.txt")
    rc, info, cut, beta = read(capi, path)
    assert rc == 0 and info.legacy == 1
    assert (info.n_species, info.n_max, info.l_max, info.beta_size) == (2, 3, 3, 3570)
    assert info.n_descriptors == 84 and info.cutoff_name == b"cosine"
    vals = tokens_after(path, 4)
    np.testing.assert_array_equal(beta, vals[1:1 + 7140])
    np.testing.assert_array_equal(cut, [5.0] * 4)


@pytest.mark.parametrize("power", [1, 2])
def test_write_read_round_trip(capi, tmp_path, power):
    from flare_pp_amd.pair import write_coefficients
    S, N, L = 2, 3, 2
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    bs = F if power == 1 else F * (F + 1) // 2
    rng = np.random.default_rng(power)
    beta = rng.normal(0, 1, (S, bs)) * 10.0 ** rng.integers(-8, 3, (S, bs))
    beta[0, 0] = 0.0
    cut = np.array([[5.0, 4.5], [4.5, 4.123]])
    p = tmp_path / "pot.txt"
    write_coefficients(p, "tester", power, S, N, L, "chebyshev", "quadratic", cut, beta)
    rc, info, cut2, beta2 = read(capi, p)
    assert rc == 0 and info.legacy == 0 and info.power == power
    np.testing.assert_array_equal(beta2, beta.ravel())  # %.16e round-trips
    np.testing.assert_array_equal(cut2, np.round(cut, 2).ravel())  # 2 decimals
    lines = p.read_text().splitlines()
    assert lines[0].startswith("DATE: ") and "CONTRIBUTOR: tester" in lines[0]
    assert lines[1] == str(power) and lines[2] == "chebyshev"
    assert lines[3] == f"{S} {N} {L} {bs}" and lines[4] == "quadratic"
    assert lines[5].split() == ["5.00", "4.50", "4.50", "4.12"]
    # five values a line; padding two spaces for positives, one otherwise
    assert all(len(line.split()) <= 5 for line in lines[6:])
    assert lines[6].startswith(" 0.0000000000000000e+00")


def test_malformed_files(capi, tmp_path):
    bad = tmp_path / "short.txt"
    bad.write_text("DATE\n2\nchebyshev\n1 2 1 30\nquadratic\n5.00\n1.0 2.0\n")
    assert read(capi, bad)[0] == 5  # EFORMAT: values missing
    size = tmp_path / "size.txt"
    size.write_text("DATE\n2\nchebyshev\n1 2 1 29\nquadratic\n5.00\n")
    assert read(capi, size)[0] == 5  # beta size mismatch
    p3 = tmp_path / "p3.txt"
    p3.write_text("DATE\n3\nchebyshev\n1 2 1 6\nquadratic\n5.00\n")
    assert read(capi, p3)[0] == 1  # power must be 1 or 2
    assert read(capi, tmp_path / "missing.txt")[0] == 4  # EIO
    msg = capi.lib().flare_last_error().decode()
    assert "Cannot open" in msg
