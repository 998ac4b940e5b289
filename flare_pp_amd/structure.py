"""Structure, B2 and the descriptor containers, mirroring flare++'s
``_C_flare.Structure`` / ``B2`` / ``DescriptorValues`` / ``ClusterDescriptor``
(ace_binding.cpp:28-106) on top of libflare_hip.

The neighbor list (structure.cpp:52-118) and every descriptor are computed on
the GPU; attributes are returned as numpy arrays with the reference's shapes
and names.
"""

import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import B2Params, check, dptr, lib

_hip = None


def _hip_lib():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_int]
    return _hip


def d2h(ptr, count, dtype):
    """Copy `count` elements of `dtype` at device address `ptr` to numpy."""
    out = np.empty(count, dtype=dtype)
    if count:
        rc = _hip_lib().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy device->host failed ({rc})")
    return out


def to_device(a, dtype, device=0):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype,
                           device=f"cuda:{device}")


class B2:
    """B2 power spectrum descriptor (b2.h:12-53).

    B2(radial_basis, cutoff_function, radial_hyps, cutoff_hyps,
       descriptor_settings[, cutoffs]) with descriptor_settings =
    [n_species, n_max, l_max].  Only the chebyshev basis with the quadratic
    or cosine cutoff is implemented (the path's settings)."""

    descriptor_name = "B2"

    def __init__(self, radial_basis, cutoff_function, radial_hyps, cutoff_hyps,
                 descriptor_settings, cutoffs=None):
        if radial_basis != "chebyshev":
            raise ValueError(f"unsupported radial basis {radial_basis!r}")
        if cutoff_function not in _capi.CUTOFF_CODES:
            raise ValueError(f"unsupported cutoff function {cutoff_function!r}")
        self.radial_basis = radial_basis
        self.cutoff_function = cutoff_function
        self.radial_hyps = [float(v) for v in radial_hyps]
        self.cutoff_hyps = [float(v) for v in cutoff_hyps]
        self.descriptor_settings = [int(v) for v in descriptor_settings]
        S = self.descriptor_settings[0]
        if cutoffs is None:
            # b2.cpp:29-31: every pair gets radial_hyps[1]
            cutoffs = np.full((S, S), self.radial_hyps[1])
        self.cutoffs = np.ascontiguousarray(np.asarray(cutoffs, dtype=np.float64)
                                            .reshape(S, S))
        self._cut_c = _capi.dbl_array(self.cutoffs.ravel())

    @property
    def n_species(self):
        return self.descriptor_settings[0]

    @property
    def n_max(self):
        return self.descriptor_settings[1]

    @property
    def l_max(self):
        return self.descriptor_settings[2]

    @property
    def n_descriptors(self):
        nr = self.n_species * self.n_max
        return nr * (nr + 1) // 2 * (self.l_max + 1)

    def params(self):
        p = B2Params()
        p.n_species, p.n_max, p.l_max = self.descriptor_settings
        p.radial_basis = _capi.RADIAL_CHEBYSHEV
        p.cutoff_function = _capi.CUTOFF_CODES[self.cutoff_function]
        p.radial_r1 = self.radial_hyps[0]
        p.cutoffs = ctypes.cast(self._cut_c, ctypes.POINTER(ctypes.c_double))
        return p

    def compute_struc(self, structure):
        return DescriptorValues._compute(self, structure)


class DescriptorValues:
    """Descriptors of one structure (descriptor.h:31-57): per-species lists.

    The force-derivative fields (descriptor_force_dervs, descriptor_force_dots,
    neighbor_coordinates, neighbor_counts, cumulative_neighbor_counts,
    neighbor_indices) are computed on first access, since the per-neighbor
    tensor is only needed by the reference's materialised kernel path."""

    _LAZY = ("descriptor_force_dervs", "descriptor_force_dots",
             "neighbor_coordinates", "neighbor_counts",
             "cumulative_neighbor_counts", "neighbor_indices", "cutoff_dervs",
             "n_neighbors_by_type")

    def __init__(self):
        self.n_descriptors = 0
        self.n_types = 0
        self.n_atoms = 0
        self.volume = 0.0
        self.descriptors = []
        self.descriptor_norms = []
        self.cutoff_values = []
        self.atom_indices = []
        self.n_clusters = 0
        self.n_clusters_by_type = []
        self.cumulative_type_count = [0]
        self._dervs = None
        self._calc = None
        self._struc = None

    @classmethod
    def _compute(cls, calc, struc):
        ctx = _capi.context(struc.device)
        p = calc.params()
        n, F, S = struc.noa, calc.n_descriptors, calc.n_species
        dev = f"cuda:{struc.device}"
        d_desc = torch.empty((max(n, 1), F), dtype=torch.float64, device=dev)
        d_norm = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        check(lib().flare_b2_compute(ctx.handle, ctypes.byref(p),
                                     struc._nlist, dptr(struc.d_species),
                                     dptr(d_desc), dptr(d_norm), None))
        desc = d_desc[:n].cpu().numpy()
        norms = d_norm[:n].cpu().numpy()
        v = cls()
        v._calc, v._struc = calc, struc
        v.n_descriptors = F
        v.n_types = S
        v.n_atoms = n
        v.volume = struc.volume
        counts = struc.species_count[:S]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(int)
        atom_rows = struc._atom_of_row()
        for s in range(S):
            a, b = off[s], off[s + 1]
            v.descriptors.append(desc[a:b])
            v.descriptor_norms.append(norms[a:b])
            v.cutoff_values.append(np.ones(b - a))
            v.atom_indices.append(atom_rows[a:b].astype(np.int32))
            v.n_clusters_by_type.append(int(b - a))
            v.cumulative_type_count.append(int(b))
        v.n_clusters = int(n)
        return v

    def _compute_dervs(self):
        struc, calc = self._struc, self._calc
        ctx = _capi.context(struc.device)
        p = calc.params()
        S, F, n = calc.n_species, calc.n_descriptors, struc.noa
        npairs = ctypes.c_int64()
        by_type = (ctypes.c_int64 * S)()
        check(lib().flare_b2_count_pairs(ctx.handle, ctypes.byref(p), struc._nlist,
                                         dptr(struc.d_species), ctypes.byref(npairs),
                                         by_type))
        P = npairs.value
        dev = f"cuda:{struc.device}"
        fd = torch.empty((max(3 * P, 1), F), dtype=torch.float64, device=dev)
        fdots = torch.empty(max(3 * P, 1), dtype=torch.float64, device=dev)
        co = torch.empty((max(P, 1), 3), dtype=torch.float64, device=dev)
        ncount = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        ncum = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        nidx = torch.empty(max(P, 1), dtype=torch.int32, device=dev)
        check(lib().flare_b2_compute_dervs(ctx.handle, ctypes.byref(p), struc._nlist,
                                           dptr(struc.d_species), dptr(fd),
                                           dptr(fdots), dptr(co), dptr(ncount),
                                           dptr(ncum), dptr(nidx)))
        fd, fdots, co = fd[:3 * P].cpu().numpy(), fdots[:3 * P].cpu().numpy(), \
            co[:P].cpu().numpy()
        ncount, ncum, nidx = ncount[:n].cpu().numpy(), ncum[:n].cpu().numpy(), \
            nidx[:P].cpu().numpy()
        out = {k: [] for k in self._LAZY}
        poff = np.concatenate([[0], np.cumsum(list(by_type))]).astype(int)
        roff = self.cumulative_type_count
        for s in range(S):
            a, b = poff[s], poff[s + 1]
            out["descriptor_force_dervs"].append(fd[3 * a:3 * b])
            out["descriptor_force_dots"].append(fdots[3 * a:3 * b])
            out["neighbor_coordinates"].append(co[a:b])
            out["neighbor_indices"].append(nidx[a:b])
            out["cutoff_dervs"].append(np.zeros(3 * (b - a)))
            out["neighbor_counts"].append(ncount[roff[s]:roff[s + 1]])
            out["cumulative_neighbor_counts"].append(ncum[roff[s]:roff[s + 1]])
            out["n_neighbors_by_type"].append(int(b - a))
        self._dervs = out

    def __getattr__(self, name):
        if name in DescriptorValues._LAZY:
            if self.__dict__.get("_dervs") is None:
                self._compute_dervs()
            return self._dervs[name]
        raise AttributeError(name)


class ClusterDescriptor:
    """A set of environments (descriptor.h:61-90; descriptor.cpp:62-176)."""

    def __init__(self, structure=None, clusters=None):
        self.descriptors = []
        self.descriptor_norms = []
        self.cutoff_values = []
        self.n_clusters_by_type = []
        self.cumulative_type_count = []
        self.n_descriptors = 0
        self.n_types = 0
        self.n_clusters = 0
        if structure is not None:
            if clusters is None:
                self.add_all_clusters(structure)
            elif len(clusters) and isinstance(clusters[0], (list, tuple, np.ndarray)):
                self.add_clusters_by_type(structure, clusters)
            else:
                self.add_clusters(structure, clusters)

    def initialize_cluster(self, n_types, n_descriptors):
        if self.n_clusters != 0:
            return
        self.n_types = n_types
        self.n_descriptors = n_descriptors
        self.descriptors = [np.zeros((0, n_descriptors)) for _ in range(n_types)]
        self.descriptor_norms = [np.zeros(0) for _ in range(n_types)]
        self.cutoff_values = [np.zeros(0) for _ in range(n_types)]
        self.n_clusters_by_type = [0] * n_types
        self.cumulative_type_count = [0] * n_types

    def add_clusters_by_type(self, structure, clusters):
        self.initialize_cluster(structure.n_types, structure.n_descriptors)
        for s in range(self.n_types):
            idx = np.asarray(clusters[s], dtype=int)
            self.descriptors[s] = np.vstack([self.descriptors[s],
                                             structure.descriptors[s][idx]])
            self.descriptor_norms[s] = np.concatenate(
                [self.descriptor_norms[s], structure.descriptor_norms[s][idx]])
            self.cutoff_values[s] = np.concatenate(
                [self.cutoff_values[s], structure.cutoff_values[s][idx]])
        for s in range(self.n_types):
            self.n_clusters_by_type[s] += len(clusters[s])
            self.n_clusters += len(clusters[s])
            if s > 0:
                self.cumulative_type_count[s] = (self.cumulative_type_count[s - 1]
                                                 + self.n_clusters_by_type[s - 1])

    def add_clusters(self, structure, clusters):
        by_type = [[] for _ in range(structure.n_types)]
        ctc = structure.cumulative_type_count
        for c in clusters:
            for s in range(structure.n_types):
                if ctc[s] <= c < ctc[s + 1]:
                    by_type[s].append(c - ctc[s])
                    break
        self.add_clusters_by_type(structure, by_type)

    def add_all_clusters(self, structure):
        self.add_clusters_by_type(
            structure, [list(range(k)) for k in structure.n_clusters_by_type])

    def stacked(self):
        """(M x F) species-grouped matrix and per-species counts."""
        F = self.n_descriptors
        mats = [d.reshape(-1, F) for d in self.descriptors]
        return (np.ascontiguousarray(np.vstack(mats)) if mats else np.zeros((0, F)),
                list(self.n_clusters_by_type))


class Structure:
    """Periodic structure with its neighbor list and descriptors
    (structure.h:8-117).  Structure(cell, species, positions[, cutoff,
    descriptor_calculators]); cell rows are lattice vectors, species are
    0-based codes."""

    def __init__(self, cell, species, positions, cutoff=None,
                 descriptor_calculators=None, device=0):
        self.device = device
        self.cell = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3))
        self.species = [int(s) for s in species]
        self.positions = np.ascontiguousarray(np.asarray(positions, dtype=np.float64)
                                              .reshape(-1, 3))
        self.noa = len(self.species)
        if self.positions.shape[0] != self.noa:
            raise ValueError("positions and species lengths differ")
        self.cell_transpose = self.cell.T.copy()
        self.energy = np.zeros(0)
        self.forces = np.zeros(0)
        self.stresses = np.zeros(0)
        self.mean_efs = None
        self.variance_efs = None
        self.local_uncertainties = []
        self.descriptors = []
        self.descriptor_calculators = list(descriptor_calculators or [])
        self.d_positions = to_device(self.positions, torch.float64, device)
        self.d_species = to_device(np.asarray(self.species, dtype=np.int32),
                                   torch.int32, device)
        self._nlist = None
        self._view = None
        self.cutoff = None
        if cutoff is not None:
            self.cutoff = float(cutoff)
            self.compute_neighbors()
            self.compute_descriptors()

    # -- neighbor list --------------------------------------------------
    def compute_neighbors(self):
        ctx = _capi.context(self.device)
        if self._nlist is None:
            h = ctypes.c_void_p()
            check(lib().flare_nlist_create(ctx.handle, ctypes.byref(h)))
            self._nlist = h
        cell = _capi.dbl_array(self.cell.ravel())
        check(lib().flare_nlist_build(ctx.handle, self._nlist, self.noa, cell,
                                      dptr(self.d_positions), dptr(self.d_species),
                                      self.cutoff))
        v = _capi.NListView()
        check(lib().flare_nlist_get_view(self._nlist, ctypes.byref(v)))
        self._view = v
        self._cache = {}

    @property
    def volume(self):
        return self._view.volume if self._view is not None else \
            abs(np.linalg.det(self.cell))

    @property
    def n_neighbors(self):
        return int(self._view.n_pairs)

    @property
    def sweep(self):
        return int(self._view.sweep)

    @property
    def single_sweep_cutoff(self):
        return float(self._view.single_sweep_cutoff)

    @property
    def species_count(self):
        return np.array(list(self._view.species_count), dtype=int)

    def _get(self, key, ptr, count, dtype, shape=None):
        if key not in self._cache:
            a = d2h(ptr, count, dtype)
            self._cache[key] = a.reshape(shape) if shape else a
        return self._cache[key]

    @property
    def neighbor_count(self):
        return self._get("nc", self._view.d_neighbor_count, self.noa, np.int32)

    @property
    def cumulative_neighbor_count(self):
        return self._get("cnc", self._view.d_cumulative_neighbor_count,
                         self.noa + 1, np.int32)

    @property
    def structure_indices(self):
        return self._get("si", self._view.d_structure_indices, self.n_neighbors,
                         np.int32)

    @property
    def neighbor_species(self):
        return self._get("ns", self._view.d_neighbor_species, self.n_neighbors,
                         np.int32)

    @property
    def relative_positions(self):
        return self._get("rp", self._view.d_relative_positions,
                         4 * self.n_neighbors, np.float64, (self.n_neighbors, 4))

    @property
    def wrapped_positions(self):
        return self._get("wp", self._view.d_wrapped_positions, 3 * self.noa,
                         np.float64, (self.noa, 3))

    def wrap_positions(self):
        return self.wrapped_positions

    def _atom_of_row(self):
        return self._get("aor", self._view.d_atom_indices, self.noa, np.int32)

    # -- descriptors ------------------------------------------------------
    def compute_descriptors(self):
        self.descriptors = [calc.compute_struc(self)
                            for calc in self.descriptor_calculators]

    def __del__(self):
        h = getattr(self, "_nlist", None)
        if h and _capi._lib is not None:
            _capi._lib.flare_nlist_destroy(h)
            self._nlist = None
