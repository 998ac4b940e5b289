"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py.  The product package never imports it.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

CUTOFF_QUADRATIC = 0
CUTOFF_COSINE = 1
_CUTOFF_CODES = {"quadratic": CUTOFF_QUADRATIC, "cosine": CUTOFF_COSINE}

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build():
    """Compile liboracle.so with the committed Makefile (host g++ only)."""
    subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


def build_ref():
    """Compile the reference's own radial/cutoff sources into _ref/ when the
    flare_pp source tree is present; returns True when built."""
    src = "/root/reference/src/flare_pp/radial.cpp"
    if not os.path.exists(src):
        return False
    subprocess.run(["make", "-C", _HERE, "ref"], check=True,
                   stdout=subprocess.DEVNULL)
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_single_sweep_cutoff": (ctypes.c_double, [_d]),
            "orc_volume": (ctypes.c_double, [_d]),
            "orc_inverse3": (None, [_d, _d]),
            "orc_wrap_positions": (None, [_d, ctypes.c_int, _d, _d]),
            "orc_nlist_brute": (_vp, [_d, ctypes.c_int, _d, _i, ctypes.c_double]),
            "orc_nlist_cells": (_vp, [_d, ctypes.c_int, _d, _i, ctypes.c_double,
                                      ctypes.c_int]),
            "orc_nlist_brute_subset": (_vp, [_d, ctypes.c_int, _d, _i,
                                             ctypes.c_double, ctypes.c_int, _i]),
            "orc_nlist_free": (None, [_vp]),
            "orc_nlist_n_pairs": (ctypes.c_int64, [_vp]),
            "orc_nlist_n_atoms": (ctypes.c_int, [_vp]),
            "orc_nlist_sweep": (ctypes.c_int, [_vp]),
            "orc_nlist_get": (None, [_vp, _i, _i, _i, _i, _d, _i, _d]),
            "orc_radial": (None, [ctypes.c_double] * 5 + [ctypes.c_int,
                                                         ctypes.c_double,
                                                         ctypes.c_int, _d, _d,
                                                         _d, _d]),
            "orc_ylm": (None, [ctypes.c_double] * 3 + [ctypes.c_int, _d, _d, _d, _d]),
            "orc_b2_compute": (_vp, [_vp, _i, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     _d, ctypes.c_int]),
            "orc_b2_free": (None, [_vp]),
            "orc_b2_n_descriptors": (ctypes.c_int, [_vp]),
            "orc_b2_n_pairs": (ctypes.c_int64, [_vp]),
            "orc_b2_get_atom_order": (None, [_vp, _d, _d, _d]),
            "orc_b2_get_grouped": (None, [_vp, _i, _i, _d, _d, _d, _d, _d, _i,
                                          _i, _i, _i]),
            "orc_b2_get_single_bond_dervs": (None, [_vp, _d]),
            "orc_envs_envs": (None, [ctypes.c_int, ctypes.c_int, _i, _d, _d, _i,
                                     _d, _d, ctypes.c_double, ctypes.c_double, _d]),
            "orc_envs_struc": (None, [_vp, _i, _d, _d, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _d]),
            "orc_cholesky_inverse": (ctypes.c_int, [ctypes.c_int, _d,
                                                    ctypes.c_double, _d]),
            "orc_predict_mean": (None, [_vp, _i, _d, _d, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, _d, _d]),
            "orc_cluster_uncertainties": (None, [_vp, _i, _d, _d,
                                                 ctypes.c_double,
                                                 ctypes.c_double, _d, _d]),
            "orc_cluster_uncertainties_rows": (None, [_vp, _i, _d, _d,
                                                      ctypes.c_double,
                                                      ctypes.c_double, _d,
                                                      ctypes.c_int, _i, _d]),
            "orc_mapping_coefficients": (ctypes.c_int, [ctypes.c_int,
                                                        ctypes.c_int, _i, _d, _d,
                                                        ctypes.c_double,
                                                        ctypes.c_double, _d, _d]),
            "orc_pair_compute": (None, [_vp, _i, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _d,
                                        ctypes.c_int, _d, ctypes.c_int, _d, _d,
                                        _d, _d]),
            "orc_pair_compute_full": (None, [_vp, _i, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, _d,
                                             ctypes.c_int, _d, ctypes.c_int, _d,
                                             _d, _d, _d]),
            "orc_unpack_beta": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _d, _d]),
            "orc_pair_compute_lmp": (None, [ctypes.c_int, _d, _i, ctypes.c_int, _i,
                                            _i, _i, _i, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, _d,
                                            ctypes.c_int, _d, _d, _d, _d, _d]),
            "orc_self_kernel_struc": (None, [_vp, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, _d]),
            "orc_varmap_coefficients": (ctypes.c_int, [ctypes.c_int, ctypes.c_int,
                                                       _i, _d, _d,
                                                       ctypes.c_double, _d, _d]),
            "orc_std_atom": (None, [_vp, _i, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _d, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int, _d,
                                    ctypes.c_int, ctypes.c_int, _i, _d, _d,
                                    ctypes.c_int, _d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_d) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(_i) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def cutoff_code(name):
    return _CUTOFF_CODES[name]


# ---------------------------------------------------------------- geometry
def volume(cell):
    return lib().orc_volume(_pd(_f64(cell).ravel()))


def single_sweep_cutoff(cell):
    return lib().orc_single_sweep_cutoff(_pd(_f64(cell).ravel()))


def inverse3(a):
    out = np.zeros(9)
    lib().orc_inverse3(_pd(_f64(a).ravel()), _pd(out))
    return out.reshape(3, 3)


def wrap_positions(cell, positions):
    p = _f64(positions)
    out = np.zeros_like(p)
    lib().orc_wrap_positions(_pd(_f64(cell).ravel()), p.shape[0], _pd(p), _pd(out))
    return out


class NeighborList:
    """Structure neighbor list (structure.cpp:52-118) held by the oracle."""

    def __init__(self, cell, species, positions, cutoff, method="cells",
                 subset=None, nthreads=0):
        self.cell = _f64(cell).reshape(3, 3)
        self.species = _i32(species)
        self.positions = _f64(positions)
        self.cutoff = float(cutoff)
        n = self.positions.shape[0]
        L = lib()
        c = _pd(self.cell.ravel())
        if subset is not None:
            sub = _i32(subset)
            self._h = L.orc_nlist_brute_subset(c, n, _pd(self.positions),
                                               _pi(self.species), self.cutoff,
                                               len(sub), _pi(sub))
            rows = len(sub)
        elif method == "brute":
            self._h = L.orc_nlist_brute(c, n, _pd(self.positions),
                                        _pi(self.species), self.cutoff)
            rows = n
        else:
            self._h = L.orc_nlist_cells(c, n, _pd(self.positions),
                                        _pi(self.species), self.cutoff, nthreads)
            rows = n
        self._free = L.orc_nlist_free
        P = L.orc_nlist_n_pairs(self._h)
        self.n_rows = rows
        self.count = np.zeros(rows, np.int32)
        self.cum = np.zeros(rows + 1, np.int32)
        self.idx = np.zeros(P, np.int32)
        self.nspec = np.zeros(P, np.int32)
        self.rel = np.zeros((P, 4))
        self.shift = np.zeros((P, 3), np.int32)
        self.wrapped = np.zeros((n, 3))
        L.orc_nlist_get(self._h, _pi(self.count), _pi(self.cum), _pi(self.idx),
                        _pi(self.nspec), _pd(self.rel), _pi(self.shift),
                        _pd(self.wrapped))
        self.sweep = L.orc_nlist_sweep(self._h)
        self.n_pairs = int(P)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        # the library handle is looked up at construction: module globals may
        # already be gone when the interpreter shuts down
        h = getattr(self, "_h", None)
        free = getattr(self, "_free", None)
        if h and free is not None:
            free(h)
            self._h = None


# ------------------------------------------------------- radial, harmonics
def radial(x, y, z, r, rcut, N, r1=0.0, cutoff="quadratic"):
    g, gx, gy, gz = (np.zeros(N) for _ in range(4))
    lib().orc_radial(x, y, z, r, rcut, N, r1, cutoff_code(cutoff), _pd(g),
                     _pd(gx), _pd(gy), _pd(gz))
    return g, gx, gy, gz


def ylm(x, y, z, L):
    H = (L + 1) ** 2
    out = [np.zeros(H) for _ in range(4)]
    lib().orc_ylm(x, y, z, L, *[_pd(a) for a in out])
    return out


# ------------------------------------------------------------------- B2
class B2Result:
    """B2 descriptor values (b2.cpp:76-399) held by the oracle."""

    def __init__(self, nlist, species, S, N, L, cutoffs, r1=0.0,
                 cutoff="quadratic", dervs=False):
        self.nlist = nlist  # keep alive
        self.S, self.N, self.L = S, N, L
        self.cutoffs = _f64(cutoffs).reshape(S, S)
        self.species = _i32(species)
        self._h = lib().orc_b2_compute(nlist.handle, _pi(self.species), S, N, L,
                                       r1, cutoff_code(cutoff),
                                       _pd(self.cutoffs.ravel()), int(dervs))
        self._free = lib().orc_b2_free
        self.dervs = dervs
        self.F = lib().orc_b2_n_descriptors(self._h)
        self.H = (L + 1) ** 2
        self.W = S * N * self.H
        self.n = len(self.species)
        self.n_pairs = int(lib().orc_b2_n_pairs(self._h))

    @property
    def handle(self):
        return self._h

    def atom_order(self):
        A = np.zeros((self.n, self.W))
        B = np.zeros((self.n, self.F))
        nrm = np.zeros(self.n)
        lib().orc_b2_get_atom_order(self._h, _pd(A), _pd(B), _pd(nrm))
        return A, B, nrm

    def single_bond_dervs(self):
        d = np.zeros((3 * self.n_pairs, self.W))
        lib().orc_b2_get_single_bond_dervs(self._h, _pd(d))
        return d

    def grouped(self):
        n, F, P = self.n, self.F, self.n_pairs
        out = dict(
            n_by_type=np.zeros(self.S, np.int32),
            neigh_by_type=np.zeros(self.S, np.int32),
            descriptors=np.zeros((n, F)),
            norms=np.zeros(n),
            force_dervs=np.zeros((3 * P, F)) if self.dervs else None,
            force_dots=np.zeros(3 * P) if self.dervs else None,
            neighbor_coords=np.zeros((P, 3)),
            neighbor_counts=np.zeros(n, np.int32),
            cum_neighbor_counts=np.zeros(n, np.int32),
            atom_indices=np.zeros(n, np.int32),
            neighbor_indices=np.zeros(P, np.int32),
        )
        o = out
        lib().orc_b2_get_grouped(
            self._h, _pi(o["n_by_type"]), _pi(o["neigh_by_type"]),
            _pd(o["descriptors"]), _pd(o["norms"]), _pd(o["force_dervs"]),
            _pd(o["force_dots"]), _pd(o["neighbor_coords"]),
            _pi(o["neighbor_counts"]), _pi(o["cum_neighbor_counts"]),
            _pi(o["atom_indices"]), _pi(o["neighbor_indices"]))
        return out

    def __del__(self):
        # the library handle is looked up at construction: module globals may
        # already be gone when the interpreter shuts down
        h = getattr(self, "_h", None)
        free = getattr(self, "_free", None)
        if h and free is not None:
            free(h)
            self._h = None


# --------------------------------------------------------- kernels and GP
def envs_envs(S, n1_by_type, d1, norms1, n2_by_type, d2, norms2, sigma, power):
    d1, d2 = _f64(d1), _f64(d2)
    n1, n2 = d1.shape[0], d2.shape[0]
    out = np.zeros((n1, n2))
    lib().orc_envs_envs(S, d1.shape[1], _pi(_i32(n1_by_type)), _pd(d1),
                        _pd(_f64(norms1)), _pi(_i32(n2_by_type)), _pd(d2),
                        _pd(_f64(norms2)), sigma, power, _pd(out))
    return out


def envs_struc(b2, m_by_type, sparse, sparse_norms, sigma, power, volume_):
    sparse = _f64(sparse)
    out = np.zeros((sparse.shape[0], 1 + 3 * b2.n + 6))
    lib().orc_envs_struc(b2.handle, _pi(_i32(m_by_type)), _pd(sparse),
                         _pd(_f64(sparse_norms)), sigma, power, volume_, _pd(out))
    return out


def cholesky_inverse(K, jitter):
    K = _f64(K)
    out = np.zeros_like(K)
    rc = lib().orc_cholesky_inverse(K.shape[0], _pd(K), jitter, _pd(out))
    if rc:
        raise np.linalg.LinAlgError("oracle Cholesky failed")
    return out


def predict_mean(b2, m_by_type, sparse, sparse_norms, sigma, power, volume_,
                 alpha):
    out = np.zeros(1 + 3 * b2.n + 6)
    lib().orc_predict_mean(b2.handle, _pi(_i32(m_by_type)), _pd(_f64(sparse)),
                           _pd(_f64(sparse_norms)), sigma, power, volume_,
                           _pd(_f64(alpha)), _pd(out))
    return out


def cluster_uncertainties(b2, m_by_type, sparse, sparse_norms, sigma, power,
                          Linv, rows=None):
    """Local variances in cluster (species-grouped) order; with rows, only
    those grouped rows (ascending), in that order."""
    if rows is None:
        out = np.zeros(b2.n)
        lib().orc_cluster_uncertainties(b2.handle, _pi(_i32(m_by_type)),
                                        _pd(_f64(sparse)), _pd(_f64(sparse_norms)),
                                        sigma, power, _pd(_f64(Linv)), _pd(out))
        return out
    rows = _i32(np.sort(np.asarray(rows)))
    out = np.zeros(len(rows))
    lib().orc_cluster_uncertainties_rows(b2.handle, _pi(_i32(m_by_type)),
                                         _pd(_f64(sparse)), _pd(_f64(sparse_norms)),
                                         sigma, power, _pd(_f64(Linv)), len(rows),
                                         _pi(rows), _pd(out))
    return out


def mapping_coefficients(S, m_by_type, sparse, sparse_norms, sigma, power,
                         alpha):
    sparse = _f64(sparse)
    F = sparse.shape[1]
    bs = F if int(power) == 1 else F * (F + 1) // 2
    out = np.zeros((S, bs))
    rc = lib().orc_mapping_coefficients(S, F, _pi(_i32(m_by_type)), _pd(sparse),
                                        _pd(_f64(sparse_norms)), sigma, power,
                                        _pd(_f64(alpha)), _pd(out))
    if rc:
        raise ValueError("mapping defined for power 1 and 2 only")
    return out


def unpack_beta(S, F, power, packed):
    packed = _f64(packed)
    out = np.zeros((S, F) if power == 1 else (S, F, F))
    lib().orc_unpack_beta(S, F, power, _pd(packed), _pd(out))
    return out


def pair_compute(nlist, species, S, N, L, power, cutoffs, beta, cutoff="quadratic",
                 packed=True, nthreads=0, want_eatom=False):
    """Mapped pair_flare energy, forces and virial (xx yy zz xy xz yz)."""
    n = len(species)
    e = np.zeros(1)
    f = np.zeros((n, 3))
    v = np.zeros(6)
    ea = np.zeros(n) if want_eatom else None
    fn = lib().orc_pair_compute if packed else lib().orc_pair_compute_full
    fn(nlist.handle, _pi(_i32(species)), S, N, L, power,
       _pd(_f64(cutoffs).ravel()), cutoff_code(cutoff), _pd(_f64(beta)),
       nthreads, _pd(e), _pd(f), _pd(v), _pd(ea))
    return (e[0], f, v, ea) if want_eatom else (e[0], f, v)


def varmap_coefficients(S, m_by_type, sparse, sparse_norms, sigma, Kinv):
    """Variance map S x F x F of a power-1 kernel
    (normalized_dot_product.cpp:761-852)."""
    sparse = _f64(sparse)
    F = sparse.shape[1]
    out = np.zeros((S, F, F))
    lib().orc_varmap_coefficients(S, F, _pi(_i32(m_by_type)), _pd(sparse),
                                  _pd(_f64(sparse_norms)), sigma,
                                  _pd(_f64(Kinv)), _pd(out))
    return out


def std_atom_map(nlist, species, S, N, L, cutoffs, sigma, beta_var,
                 cutoff="quadratic", nthreads=0):
    """compute flare/std/atom with a variance-map file
    (compute_flare_std_atom.cpp:186-191)."""
    n = len(species)
    out = np.zeros(n)
    lib().orc_std_atom(nlist.handle, _pi(_i32(species)), S, N, L,
                       _pd(_f64(cutoffs).ravel()), cutoff_code(cutoff), sigma, 1,
                       _pd(_f64(beta_var)), 0, 0, None, None, None, nthreads,
                       _pd(out))
    return out


def std_atom_linv(nlist, species, S, N, L, cutoffs, sigma, power, n_by_type,
                  sparse_normed, Linv, cutoff="quadratic", nthreads=0):
    """compute flare/std/atom with L^-1 and sparse-descriptor files
    (compute_flare_std_atom.cpp:192-206)."""
    n = len(species)
    out = np.zeros(n)
    sp = _f64(sparse_normed)
    lib().orc_std_atom(nlist.handle, _pi(_i32(species)), S, N, L,
                       _pd(_f64(cutoffs).ravel()), cutoff_code(cutoff), sigma, 0,
                       None, int(power), sp.shape[0], _pi(_i32(n_by_type)),
                       _pd(sp), _pd(_f64(Linv)), nthreads, _pd(out))
    return out


def pair_compute_lmp(x, type_, ilist, numneigh, offsets, neigh, S, N, L, power,
                     cutoffs, beta_full, cutoff="quadratic"):
    """PairFLARE::compute on a LAMMPS-style list with ghosts: energy, forces
    (n_all x 3, ghost rows included), virial (xx yy zz xy xz yz), eatom."""
    x = _f64(x)
    n_all = len(x)
    e = np.zeros(1)
    f = np.zeros((n_all, 3))
    v = np.zeros(6)
    ea = np.zeros(n_all)
    lib().orc_pair_compute_lmp(n_all, _pd(x), _pi(_i32(type_)), len(ilist),
                               _pi(_i32(ilist)), _pi(_i32(numneigh)),
                               _pi(_i32(offsets)), _pi(_i32(neigh)), S, N, L,
                               power, _pd(_f64(cutoffs).ravel()),
                               cutoff_code(cutoff), _pd(_f64(beta_full)), _pd(e),
                               _pd(f), _pd(v), _pd(ea))
    return e[0], f, v, ea


def self_kernel_struc(b2, sigma, power, volume_):
    """k(o, o) for o = [E, F (3n), S (6)] (normalized_dot_product.cpp:446-578);
    b2 must carry derivatives."""
    n = b2.n
    out = np.zeros(1 + 3 * n + 6)
    lib().orc_self_kernel_struc(b2.handle, sigma, power, volume_, _pd(out))
    return out
