"""Per-atom GP uncertainties of LAMMPS' ``compute flare/std/atom``
(compute_flare_std_atom.cpp) on the GPU.

``ComputeFlareStdAtom(varmap_file)`` takes a variance-map file,
``ComputeFlareStdAtom(L_inv_file, sparse_desc_file)`` the L^-1 and
sparse-descriptor files, as ``compute ID group flare/std/atom file [file]``
does (ComputeFlareStdAtom::coeff, :297-314).  ``compute`` returns the per-atom
vector of compute_peratom (:99-213) in atom order: sqrt of the normalised
variance, negative square root of its magnitude when it is negative, 0 for
empty environments.
"""

import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import check, dptr, lib


def read_std_files(path1, path2=None):
    """Host-only parse: (info dict, cutoffs S x S, values1, values2).

    values1 is the variance map (S x F x F) or L^-1 (M x M); values2 the
    normalised sparse rows (M x F) or None."""
    info = _capi.StdInfo()
    a1 = str(path1).encode()
    a2 = str(path2).encode() if path2 is not None else None
    check(lib().flare_std_read(a1, a2, ctypes.byref(info), None, None, None))
    S, F = info.n_species, info.n_descriptors
    cut = np.zeros((S, S))
    pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if info.use_map:
        v1 = np.zeros((S, F, F))
        v2 = None
        check(lib().flare_std_read(a1, a2, ctypes.byref(info), pd(cut), pd(v1), None))
    else:
        M = info.n_clusters
        v1 = np.zeros((max(M, 1), max(M, 1)))
        v2 = np.zeros((max(M, 1), F))
        check(lib().flare_std_read(a1, a2, ctypes.byref(info), pd(cut), pd(v1),
                                   pd(v2)))
        v1, v2 = v1[:M, :M], v2[:M]
    return _info_dict(info), cut, v1, v2


def _info_dict(info):
    d = {k: getattr(info, k) for k, _ in info._fields_}
    d["hyps"] = list(info.hyps)[:max(info.n_hyps, 0)]
    d["n_clusters_by_type"] = list(info.n_clusters_by_type)[:info.n_species]
    d["radial_basis"] = info.radial_basis.decode()
    d["cutoff_name"] = info.cutoff_name.decode()
    return d


class ComputeFlareStdAtom:
    """compute flare/std/atom over one model, evaluated on the GPU."""

    def __init__(self, path1=None, path2=None, device=0, _handle=None):
        self.device = device
        ctx = _capi.context(device)
        if _handle is None:
            h = ctypes.c_void_p()
            check(lib().flare_std_load(
                ctx.handle, str(path1).encode(),
                str(path2).encode() if path2 is not None else None,
                ctypes.byref(h)))
            _handle = h
        self._h = _handle
        info = _capi.StdInfo()
        check(lib().flare_std_get_info(self._h, ctypes.byref(info)))
        self.info = _info_dict(info)
        self.use_map = bool(info.use_map)
        self.n_species = info.n_species
        self.cutoff = info.cutoff
        self.sigma = info.hyps[0]
        self._nlist = None

    @classmethod
    def from_varmap(cls, beta_var, hyps, n_species, n_max, l_max, cutoffs,
                    cutoff_function="quadratic", device=0):
        """Variance map S x F x F (row-major) from memory."""
        info = cls._base_info(1, 2, hyps, n_species, n_max, l_max, cutoff_function)
        return cls._create(info, cutoffs, beta_var, None, device)

    @classmethod
    def from_L_inverse(cls, L_inv, sparse_normed, n_clusters_by_type, power, hyps,
                       n_species, n_max, l_max, cutoffs,
                       cutoff_function="quadratic", device=0):
        """L^-1 (M x M, lower) and normalised sparse rows (M x F) from memory."""
        info = cls._base_info(0, power, hyps, n_species, n_max, l_max,
                              cutoff_function)
        info.n_clusters = int(sum(n_clusters_by_type))
        for s, c in enumerate(n_clusters_by_type):
            info.n_clusters_by_type[s] = int(c)
        return cls._create(info, cutoffs, L_inv, sparse_normed, device)

    @staticmethod
    def _base_info(use_map, power, hyps, S, N, L, cutoff_function):
        info = _capi.StdInfo()
        info.use_map = use_map
        info.power = int(power)
        hyps = list(hyps)
        info.n_hyps = len(hyps)
        for k, h in enumerate(hyps[:4]):
            info.hyps[k] = float(h)
        info.n_species, info.n_max, info.l_max = int(S), int(N), int(L)
        info.cutoff_function = _capi.CUTOFF_CODES[cutoff_function]
        return info

    @classmethod
    def _create(cls, info, cutoffs, v1, v2, device):
        ctx = _capi.context(device)
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64))
                if a is not None else None for a in (cutoffs, v1, v2)]
        ptr = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None
               else None for a in arrs]
        h = ctypes.c_void_p()
        check(lib().flare_std_create(ctx.handle, ctypes.byref(info), *ptr,
                                     ctypes.byref(h)))
        return cls(device=device, _handle=h)

    # -- device-resident evaluation -------------------------------------------
    def build_neighbors(self, cell, d_positions, d_species):
        ctx = _capi.context(self.device)
        if self._nlist is None:
            h = ctypes.c_void_p()
            check(lib().flare_nlist_create(ctx.handle, ctypes.byref(h)))
            self._nlist = h
        cellc = _capi.dbl_array(np.asarray(cell, dtype=np.float64).ravel())
        check(lib().flare_nlist_build(ctx.handle, self._nlist, d_species.numel(),
                                      cellc, dptr(d_positions), dptr(d_species),
                                      self.cutoff))
        return self._nlist

    def compute_device(self, d_species, d_stds, nlist=None):
        ctx = _capi.context(self.device)
        nl = nlist if nlist is not None else self._nlist
        check(lib().flare_std_compute(ctx.handle, self._h, nl, dptr(d_species),
                                      dptr(d_stds)))

    def compute(self, cell, species, positions):
        """Per-atom stds (atom order) of a periodic structure."""
        dev = f"cuda:{self.device}"
        pos = torch.as_tensor(np.ascontiguousarray(positions, dtype=np.float64),
                              device=dev).reshape(-1, 3).contiguous()
        spc = torch.as_tensor(np.asarray(species, dtype=np.int32), device=dev)
        n = spc.numel()
        self.build_neighbors(cell, pos, spc)
        out = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        self.compute_device(spc, out)
        return out[:n].cpu().numpy()

    def __del__(self):
        if _capi._lib is not None:
            if getattr(self, "_nlist", None):
                _capi._lib.flare_nlist_destroy(self._nlist)
                self._nlist = None
            if getattr(self, "_h", None):
                _capi._lib.flare_std_destroy(self._h)
                self._h = None
