"""pair_style flare on the GPU: coefficient file reader/writer and the mapped
force evaluation (pair_flare.cpp:68-184, 263-416; lammps_descriptor.cpp).

``PairFLARE(path).compute(cell, species, positions)`` returns the energy,
forces and the LAMMPS virial (xx, yy, zz, xy, xz, yz) of a periodic
structure; neighbors come from the Structure list built on the GPU.
"""

import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import PairInfo, check, dptr, lib


def _form(form):
    return _capi.FORMS[form] if isinstance(form, str) else int(form)


def write_coefficients(path, contributor, power, n_species, n_max, l_max,
                       radial_basis, cutoff_function, cutoffs, beta_packed):
    """Write a coefficient file in the layout of
    SparseGP::write_mapping_coefficients (sparse_gp.cpp:1142-1206)."""
    beta = np.ascontiguousarray(np.asarray(beta_packed, dtype=np.float64))
    beta = beta.reshape(n_species, -1)
    cut = np.ascontiguousarray(np.asarray(cutoffs, dtype=np.float64).ravel())
    check(lib().flare_write_coefficients(
        str(path).encode(), str(contributor).encode(), int(power), int(n_species),
        int(n_max), int(l_max), radial_basis.encode(), cutoff_function.encode(),
        cut.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), beta.shape[1],
        beta.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


def pack_beta(beta_full):
    """Symmetric F x F -> packed upper triangle, off-diagonals doubled
    (normalized_dot_product.cpp:736-753)."""
    b = np.asarray(beta_full, dtype=np.float64)
    F = b.shape[0]
    iu = np.triu_indices(F)
    vals = b[iu].copy()
    vals[iu[0] != iu[1]] *= 2
    return vals


class PairFLARE:
    """Mapped potential loaded from a coefficient file (current or legacy
    layout) or from arrays."""

    def __init__(self, path=None, device=0, _handle=None):
        self.device = device
        ctx = _capi.context(device)
        if _handle is None:
            h = ctypes.c_void_p()
            check(lib().flare_pair_load(ctx.handle, str(path).encode(),
                                        ctypes.byref(h)))
            _handle = h
        self._h = _handle
        info = PairInfo()
        check(lib().flare_pair_get_info(self._h, ctypes.byref(info)))
        self.power = info.power
        self.n_species = info.n_species
        self.n_max = info.n_max
        self.l_max = info.l_max
        self.n_descriptors = info.n_descriptors
        self.beta_size = info.beta_size
        self.cutoff = info.cutoff
        self.legacy = bool(info.legacy)
        self.radial_basis = info.radial_basis.decode()
        self.cutoff_function = info.cutoff_name.decode()
        cut = np.zeros(self.n_species ** 2)
        check(lib().flare_pair_get_cutoffs(
            self._h, cut.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.cutoff_matrix = cut.reshape(self.n_species, self.n_species)
        self._nlist = None

    @classmethod
    def from_coefficients(cls, power, n_species, n_max, l_max, cutoffs,
                          beta_packed, cutoff_function="quadratic", device=0):
        ctx = _capi.context(device)
        cut = np.ascontiguousarray(np.asarray(cutoffs, dtype=np.float64).ravel())
        beta = np.ascontiguousarray(np.asarray(beta_packed, dtype=np.float64).ravel())
        h = ctypes.c_void_p()
        check(lib().flare_pair_create(
            ctx.handle, int(power), int(n_species), int(n_max), int(l_max),
            _capi.CUTOFF_CODES[cutoff_function],
            cut.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            beta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(h)))
        return cls(device=device, _handle=h)

    @property
    def beta(self):
        """Unpacked beta: S x F x F (power 2) or S x F (power 1)."""
        F = self.n_descriptors
        shape = (self.n_species, F) if self.power == 1 else (self.n_species, F, F)
        out = np.zeros(shape)
        ctx = _capi.context(self.device)
        check(lib().flare_pair_get_beta(
            ctx.handle, self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def beta_tensor(self):
        return torch.as_tensor(self.beta, device=f"cuda:{self.device}")

    # -- device-resident evaluation -------------------------------------------
    def build_neighbors(self, cell, d_positions, d_species):
        """Structure neighbor list at the largest pair cutoff."""
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

    def compute_device(self, d_species, form="plain", d_beta=None, energy=None,
                       forces=None, virial=None, eatom=None, nlist=None,
                       noalloc=False):
        """Force evaluation on device buffers with the last neighbor list.
        noalloc: through flare_pair_compute_noalloc, whose workspace is the
        device's static scratch (no device allocation during the call)."""
        ctx = _capi.context(self.device)
        nl = nlist if nlist is not None else self._nlist
        fn = lib().flare_pair_compute_noalloc if noalloc else lib().flare_pair_compute
        check(fn(ctx.handle, self._h, nl, dptr(d_species), _form(form), dptr(d_beta),
                 dptr(energy), dptr(forces), dptr(virial), dptr(eatom)))

    def compute_neigh_device(self, n_all, inum, d_ilist, d_numneigh, d_neighbors,
                             stride_i, stride_j, d_x, d_type, form="plain",
                             energy=None, forces=None, virial=None, eatom=None):
        """PairFLARE::compute on a caller's full neighbor list with ghost atoms
        (LAMMPS layout; forces and eatom accumulate, energy and virial are
        written)."""
        ctx = _capi.context(self.device)
        check(lib().flare_pair_compute_neigh(
            ctx.handle, self._h, int(n_all), int(inum), dptr(d_ilist),
            dptr(d_numneigh), dptr(d_neighbors), int(stride_i), int(stride_j),
            dptr(d_x), dptr(d_type), _form(form), dptr(energy), dptr(forces),
            dptr(virial), dptr(eatom)))

    def beta_grad_device(self, out):
        ctx = _capi.context(self.device)
        check(lib().flare_pair_beta_grad(ctx.handle, self._h, dptr(out)))

    # -- host convenience ---------------------------------------------------------
    def compute(self, cell, species, positions, form="plain", want_eatom=False):
        """Energy, forces (n x 3) and virial (6) of a periodic structure."""
        dev = f"cuda:{self.device}"
        pos = torch.as_tensor(np.ascontiguousarray(positions, dtype=np.float64),
                              device=dev).reshape(-1, 3).contiguous()
        spc = torch.as_tensor(np.asarray(species, dtype=np.int32), device=dev)
        n = spc.numel()
        self.build_neighbors(cell, pos, spc)
        e = torch.empty(1, dtype=torch.float64, device=dev)
        f = torch.empty((max(n, 1), 3), dtype=torch.float64, device=dev)
        v = torch.empty(6, dtype=torch.float64, device=dev)
        ea = torch.empty(max(n, 1), dtype=torch.float64, device=dev) if want_eatom else None
        self.compute_device(spc, form=form, energy=e, forces=f, virial=v, eatom=ea)
        out = (float(e.item()), f[:n].cpu().numpy(), v.cpu().numpy())
        if want_eatom:
            out = out + (ea[:n].cpu().numpy(),)
        return out

    def __del__(self):
        if _capi._lib is not None:
            if getattr(self, "_nlist", None):
                _capi._lib.flare_nlist_destroy(self._nlist)
                self._nlist = None
            if getattr(self, "_h", None):
                _capi._lib.flare_pair_destroy(self._h)
                self._h = None
