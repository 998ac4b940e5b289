"""Variance-path files on the CPU: libflare_hip's writers and parser against
the restated formats of SparseGP::write_varmap_coefficients / write_L_inverse
/ write_sparse_descriptors (sparse_gp.cpp:1208-1396) and the readers of
compute flare/std/atom (compute_flare_std_atom.cpp:349-664).  No device work
happens here."""
import ctypes

import numpy as np
import pytest

from oracle import files

DATE = "Thu Jan  1 00:00:00 2026"


@pytest.fixture(scope="module")
def capi():
    from flare_pp_amd import _capi
    _capi.lib()
    return _capi


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(
        ctypes.POINTER(ctypes.c_double))


def body(path):
    """file text without its first (DATE) line"""
    with open(path) as f:
        text = f.read()
    return text[text.index("\n") + 1:]


def expected_body(text):
    return text[text.index("\n") + 1:]


def F_of(S, N, L):
    nr = S * N
    return nr * (nr + 1) // 2 * (L + 1)


# ---- the restated writers against hand-derived text -----------------------

def test_restated_L_inverse_layout():
    # sparse_gp.cpp:1324-1337: counter starts at 1, newline after every fifth
    # value, a final newline unless the counter is a multiple of 5 -- so a
    # block of 10 values ends with an empty line
    text = files.write_L_inverse(DATE, "me", 2.0, [1.5, 0.1, 0.2, 0.3], 1, 2, 1,
                                 "chebyshev", "quadratic", [5.0],
                                 [[1, 0, 0, 0], [2, 3, 0, 0], [4, 5, 6, 0],
                                  [7, 8, 9, 10]])
    lines = text.split("\n")
    assert lines[0] == f"L_inverse_block file DATE: {DATE} CONTRIBUTOR: me"
    assert lines[1] == "2"
    assert lines[2] == "4"
    assert lines[3] == ("1.5000000000000000e+00 1.0000000000000001e-01 "
                        "2.0000000000000001e-01 2.9999999999999999e-01 ")
    assert lines[4:9] == ["chebyshev", "1 2 1 1", "quadratic", "5.00 ", "4"]
    vals = [f"{v:.16e} " for v in range(1, 11)]
    assert lines[9] == "".join(vals[:5])
    assert lines[10] == "".join(vals[5:])
    assert lines[11:] == ["", ""]  # blank line, then end of file
    three = files.write_L_inverse(DATE, "me", 1.0, [1.0], 1, 1, 0, "chebyshev",
                                  "cosine", [4.5], [[1, 0], [2, 3]])
    assert three.endswith("4.50 \n2\n" + "".join(f"{v:.16e} " for v in (1, 2, 3))
                          + "\n")


def test_restated_sparse_layout():
    # sparse_gp.cpp:1368-1391: the value counter runs on across species; a
    # species block ends with a newline unless the counter sits just past a
    # multiple of 5
    rows = [[1.0, 2.0, 2.0], [0.0, 0.0, 0.0], [3.0, 4.0, 0.0]]
    norms = [3.0, 0.0, 5.0]
    text = files.write_sparse(DATE, "me", [2, 1], rows, norms)
    lines = text.split("\n")
    assert lines[1:4] == ["1", "0 3 2", "2"]
    f = lambda v: f"{v:.16e} "
    assert lines[4] == "".join(f(v) for v in (1 / 3, 2 / 3, 2 / 3, 0.0, 0.0))
    assert lines[5] == f(0.0)
    assert lines[6] == "1"
    assert lines[7] == "".join(f(v) for v in (0.6, 0.8, 0.0))
    assert lines[8:] == [""]


def test_restated_varmap_layout():
    # sparse_gp.cpp:1243-1279: two spaces before positives, one otherwise;
    # five to a line; a new line for each species' row
    text = files.write_varmap(DATE, "me", [2.0, 0.1, 0.2, 0.3], 2, 1, 0,
                              "chebyshev", "quadratic", [5.0, 4.5, 4.5, 5.0],
                              [[1.0, -2.0, 0.0], [4.0] * 3])
    lines = text.split("\n")
    assert lines[1] == "4"
    assert lines[3:7] == ["chebyshev", "2 1 0 3", "quadratic",
                          "5.00 4.50 4.50 5.00 "]
    assert lines[7] == "  1.0000000000000000e+00 -2.0000000000000000e+00 0.0000000000000000e+00"
    assert lines[8] == "  4.0000000000000000e+00" * 3
    assert len(lines) == 9


# ---- the library's writers against the restatement ------------------------

@pytest.mark.parametrize("M", [1, 3, 4, 9, 10])
def test_write_L_inverse_matches_reference_layout(capi, tmp_path, M):
    rng = np.random.default_rng(M)
    Li = np.tril(rng.normal(size=(M, M)))
    hyps = np.array([1.7, 0.01, 0.02, 0.003])
    cut = np.array([5.0, 4.25, 4.25, 5.0])
    p = tmp_path / "Linv.txt"
    capi.check(capi.lib().flare_write_L_inverse(
        str(p).encode(), b"me", 2.0, 4, P(hyps), 2, 3, 2, b"chebyshev",
        b"quadratic", P(cut), M, P(Li)))
    want = files.write_L_inverse(DATE, "me", 2.0, hyps, 2, 3, 2, "chebyshev",
                                 "quadratic", cut, Li)
    assert body(p) == expected_body(want)
    with open(p) as f:
        head = f.readline()
    assert head.startswith("L_inverse_block file DATE: ") and head.endswith(
        " CONTRIBUTOR: me\n")


@pytest.mark.parametrize("counts", [[2, 1], [3, 0, 2], [0, 4], [5, 5]])
def test_write_sparse_descriptors_matches_reference_layout(capi, tmp_path, counts):
    rng = np.random.default_rng(len(counts))
    F = 7
    M = sum(counts)
    rows = rng.normal(size=(M, F))
    if M > 1:
        rows[1] = 0.0  # an empty environment
    norms = np.linalg.norm(rows, axis=1)
    p = tmp_path / "sparse.txt"
    arr = (ctypes.c_int * len(counts))(*counts)
    capi.check(capi.lib().flare_write_sparse_descriptors(
        str(p).encode(), b"me", len(counts), F, arr, P(rows), P(norms)))
    want = files.write_sparse(DATE, "me", counts, rows, norms)
    assert body(p) == expected_body(want)


def test_write_varmap_matches_reference_layout(capi, tmp_path):
    rng = np.random.default_rng(7)
    S, N, L = 2, 1, 1
    F = F_of(S, N, L)
    beta = rng.normal(size=(S, F * F))
    beta[0, 3] = 0.0
    hyps = np.array([2.0, 0.1, 0.2, 0.3])
    cut = np.array([5.0, 4.5, 4.5, 5.0])
    p = tmp_path / "varmap.txt"
    capi.check(capi.lib().flare_write_varmap_coefficients(
        str(p).encode(), b"me", 4, P(hyps), S, N, L, b"chebyshev", b"cosine",
        P(cut), F * F, P(beta)))
    want = files.write_varmap(DATE, "me", hyps, S, N, L, "chebyshev", "cosine",
                              cut, beta)
    assert body(p) == expected_body(want)


# ---- the library's parser -------------------------------------------------

def std_read(capi, p1, p2=None):
    info = capi.StdInfo()
    a1 = p1 and str(p1).encode()
    a2 = p2 and str(p2).encode()
    rc = capi.lib().flare_std_read(a1, a2, ctypes.byref(info), None, None, None)
    if rc:
        return rc, None, None, None, None
    S, F = info.n_species, info.n_descriptors
    cut = np.zeros(S * S)
    if info.use_map:
        v1, v2 = np.zeros(S * F * F), None
    else:
        M = info.n_clusters
        v1, v2 = np.zeros(max(M * M, 1)), np.zeros(max(M * F, 1))
    capi.check(capi.lib().flare_std_read(a1, a2, ctypes.byref(info), P(cut),
                                         P(v1) if v1 is not None else None,
                                         P(v2) if v2 is not None else None))
    return 0, info, cut, v1, v2


def test_read_varmap_written_by_restatement(capi, tmp_path):
    rng = np.random.default_rng(3)
    S, N, L = 2, 2, 1
    F = F_of(S, N, L)
    beta = rng.normal(size=(S, F * F))
    cut = [5.0, 4.5, 4.5, 3.0]
    p = tmp_path / "vm.txt"
    p.write_text(files.write_varmap(DATE, "x", [1.25, 0.1, 0.2, 0.3], S, N, L,
                                    "chebyshev", "quadratic", cut, beta))
    rc, info, c, v1, _ = std_read(capi, p)
    assert rc == 0 and info.use_map == 1 and info.power == 2
    assert (info.n_species, info.n_max, info.l_max, info.n_descriptors) == (S, N, L, F)
    assert info.beta_size == F * F and info.n_hyps == 4
    assert list(info.hyps) == [1.25, 0.1, 0.2, 0.3]
    assert info.cutoff == 5.0 and info.radial_basis == b"chebyshev"
    np.testing.assert_array_equal(c, cut)
    np.testing.assert_array_equal(v1, beta.ravel())  # %.16e round trips exactly
    ref = files.read_varmap(p.read_text())
    np.testing.assert_array_equal(np.array(ref["beta"]).ravel(), v1)


@pytest.mark.parametrize("counts", [[3, 2], [0, 4], [4, 0, 1]])
def test_read_L_inverse_and_sparse(capi, tmp_path, counts):
    rng = np.random.default_rng(11)
    S, N, L = len(counts), 1, 1
    F = F_of(S, N, L)
    M = sum(counts)
    Li = np.tril(rng.normal(size=(M, M)))
    rows = rng.normal(size=(M, F))
    norms = np.linalg.norm(rows, axis=1)
    cut = np.full(S * S, 4.0)
    p1, p2 = tmp_path / "L.txt", tmp_path / "s.txt"
    p1.write_text(files.write_L_inverse(DATE, "x", 1.0, [0.9, 0.1, 0.2, 0.3], S, N,
                                        L, "chebyshev", "cosine", cut, Li))
    p2.write_text(files.write_sparse(DATE, "x", counts, rows, norms))
    rc, info, c, v1, v2 = std_read(capi, p1, p2)
    assert rc == 0 and info.use_map == 0 and info.power == 1
    assert info.n_clusters == M and list(info.n_clusters_by_type)[:S] == counts
    assert info.cutoff_function == capi.CUTOFF_COSINE
    np.testing.assert_array_equal(v1[:M * M].reshape(M, M), Li)
    np.testing.assert_array_equal(v2[:M * F].reshape(M, F), rows / norms[:, None])


def test_reader_errors(capi, tmp_path):
    rc, *_ = std_read(capi, tmp_path / "missing.txt")
    assert capi.ERRORS[rc] == "EIO"
    S, N, L = 1, 2, 1
    F = F_of(S, N, L)
    p = tmp_path / "vm.txt"
    text = files.write_varmap(DATE, "x", [1.0], S, N, L, "chebyshev", "quadratic",
                              [5.0], [np.ones(F * F)])
    p.write_text(text.replace(f"1 2 1 {F * F}", f"1 2 1 {F * F - 1}"))
    rc, *_ = std_read(capi, p)
    assert capi.ERRORS[rc] == "EFORMAT"
    assert b"Beta size" in capi.lib().flare_last_error()
    p.write_text(text.replace("chebyshev", "bessel"))
    rc, *_ = std_read(capi, p)
    assert capi.ERRORS[rc] == "EINVAL"
    # sparse-descriptor cluster count must match the L^-1 file
    Li = np.eye(3)
    p1, p2 = tmp_path / "L.txt", tmp_path / "s.txt"
    p1.write_text(files.write_L_inverse(DATE, "x", 2.0, [1.0], 1, 1, 0, "chebyshev",
                                        "quadratic", [5.0], Li))
    p2.write_text(files.write_sparse(DATE, "x", [2], np.ones((2, 1)), np.ones(2)))
    rc, *_ = std_read(capi, p1, p2)
    assert capi.ERRORS[rc] == "EFORMAT"
    assert b"n_clusters" in capi.lib().flare_last_error()
