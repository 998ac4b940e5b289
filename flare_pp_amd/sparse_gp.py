"""NormalizedDotProduct and SparseGP prediction on the GPU, mirroring
``_C_flare.NormalizedDotProduct`` / ``_C_flare.SparseGP``
(ace_binding.cpp:150-221; normalized_dot_product.h:11-73; sparse_gp.h:13-112).

Training (likelihood, hyperparameter optimisation, QR solve for alpha) is not
part of this path: alpha is supplied with :meth:`SparseGP.set_alpha`.
"""

import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import check, dptr, lib
from .structure import ClusterDescriptor


def _form(form):
    return _capi.FORMS[form] if isinstance(form, str) else int(form)


class NormalizedDotProduct:
    """k(a, b) = sigma^2 (a.b / (|a||b|))^power (normalized_dot_product.cpp)."""

    kernel_name = "NormalizedDotProduct"

    def __init__(self, sigma, power):
        self.sigma = float(sigma)
        self.sig2 = self.sigma ** 2
        self.power = float(power)
        self.kernel_hyperparameters = np.array([self.sigma])

    def envs_envs(self, envs1, envs2, hyps=None, device=0):
        """Kernel matrix of two environment sets (normalized_dot_product.cpp:
        27-87); cross-species entries and empty environments are zero."""
        sigma = float(hyps[0]) if hyps is not None else self.sigma
        d1, n1 = envs1.stacked()
        d2, n2 = envs2.stacked()
        ctx = _capi.context(device)
        dev = f"cuda:{device}"
        t1 = torch.as_tensor(d1, device=dev)
        t2 = torch.as_tensor(d2, device=dev)
        out = torch.zeros((max(len(d1), 1), max(len(d2), 1)), dtype=torch.float64,
                          device=dev)
        S = envs1.n_types
        a1 = (ctypes.c_int * S)(*n1)
        a2 = (ctypes.c_int * S)(*n2)
        check(lib().flare_envs_envs(ctx.handle, S, envs1.n_descriptors, a1,
                                    dptr(t1), a2, dptr(t2), sigma, self.power,
                                    dptr(out)))
        return out[:len(d1), :len(d2)].cpu().numpy()

    def write_info(self, f):
        f.write(f"{int(round(self.power))}\n")


class SparseGP:
    """Sparse GP prediction (sparse_gp.cpp:612-628, 682-702) with one
    NormalizedDotProduct kernel on one B2 descriptor."""

    def __init__(self, kernels, energy_noise=0.0, force_noise=0.0,
                 stress_noise=0.0, device=0):
        if len(kernels) != 1 or not isinstance(kernels[0], NormalizedDotProduct):
            raise ValueError("one NormalizedDotProduct kernel is supported")
        self.kernels = list(kernels)
        self.n_kernels = 1
        self.energy_noise = energy_noise
        self.force_noise = force_noise
        self.stress_noise = stress_noise
        self.Kuu_jitter = 1e-8  # sparse_gp.cpp:19
        self.device = device
        self.sparse_descriptors = []
        self.sparse_indices = []
        self.descriptor_calculators = []
        self.alpha = None
        self.R_inv = None
        self._h = None
        self._dirty = True

    @property
    def n_sparse(self):
        return self.sparse_descriptors[0].n_clusters if self.sparse_descriptors else 0

    # -- sparse set (sparse_gp.cpp:56-69, 135-183) -------------------------
    def _init_sparse(self, structure):
        if not self.sparse_descriptors:
            d = structure.descriptors[0]
            c = ClusterDescriptor()
            c.initialize_cluster(d.n_types, d.n_descriptors)
            self.sparse_descriptors = [c]
            self.sparse_indices = [[]]
            self.descriptor_calculators = list(structure.descriptor_calculators)

    def add_specific_environments(self, structure, atoms):
        """Clusters whose central atom is in `atoms`, in cluster order."""
        self._init_sparse(structure)
        d = structure.descriptors[0]
        by_type = []
        for s in range(d.n_types):
            idx = []
            for k, a in enumerate(d.atom_indices[s]):
                idx.extend([k] * sum(1 for b in atoms if b == a))
            by_type.append(idx)
        self.sparse_indices[0].append(list(atoms))
        self.sparse_descriptors[0].add_clusters_by_type(d, by_type)
        self._dirty = True

    def add_all_environments(self, structure):
        self._init_sparse(structure)
        self.sparse_descriptors[0].add_all_clusters(structure.descriptors[0])
        self.sparse_indices[0].append(list(range(structure.noa)))
        self._dirty = True

    def set_alpha(self, alpha):
        alpha = np.asarray(alpha, dtype=np.float64).ravel()
        if alpha.size != self.n_sparse:
            raise ValueError("alpha length differs from the sparse set size")
        self.alpha = alpha
        if self._h is not None and not self._dirty:
            self._upload_alpha()

    # -- device model -------------------------------------------------------
    def _build(self):
        if not self._dirty and self._h is not None:
            return
        self._free()
        sp = self.sparse_descriptors[0]
        P, m_by_type = sp.stacked()
        ctx = _capi.context(self.device)
        dev = f"cuda:{self.device}"
        tP = torch.as_tensor(np.ascontiguousarray(P), device=dev)
        k = self.kernels[0]
        h = ctypes.c_void_p()
        arr = (ctypes.c_int * sp.n_types)(*m_by_type)
        check(lib().flare_sgp_create(ctx.handle, sp.n_types, sp.n_descriptors, arr,
                                     dptr(tP), k.sigma, k.power, self.Kuu_jitter,
                                     ctypes.byref(h)))
        self._h = h
        self._dirty = False
        if self.alpha is not None:
            self._upload_alpha()

    def _upload_alpha(self):
        ctx = _capi.context(self.device)
        t = torch.as_tensor(self.alpha, device=f"cuda:{self.device}")
        check(lib().flare_sgp_set_alpha(ctx.handle, self._h, dptr(t)))

    def update_L_inverse(self):
        self._build()

    def _free(self):
        if self._h is not None and _capi._lib is not None:
            _capi._lib.flare_sgp_destroy(self._h)
        self._h = None

    def __del__(self):
        self._free()

    def _matrix(self, fn):
        self._build()
        M = self.n_sparse
        ctx = _capi.context(self.device)
        out = torch.zeros((max(M, 1), max(M, 1)), dtype=torch.float64,
                          device=f"cuda:{self.device}")
        check(fn(ctx.handle, self._h, dptr(out)))
        return out[:M, :M].cpu().numpy()

    @property
    def L_inv(self):
        """L^-1 with L L^T = K_uu + jitter I (sparse_gp.cpp:581-586)."""
        return self._matrix(lib().flare_sgp_get_L_inverse)

    @property
    def Kuu(self):
        return self._matrix(lib().flare_sgp_get_Kuu)

    # -- prediction -----------------------------------------------------------
    def _check_ready(self, structure):
        if self.alpha is None:
            raise RuntimeError("alpha is not set")
        self._build()
        if not structure.descriptor_calculators:
            raise ValueError("structure has no descriptor calculator")

    def predict_mean(self, structure, form="plain"):
        """structure.mean_efs = [E, F (3n), S (6)] (sparse_gp.cpp:612-628)."""
        self._check_ready(structure)
        ctx = _capi.context(self.device)
        p = structure.descriptor_calculators[0].params()
        out = torch.empty(1 + 3 * structure.noa + 6, dtype=torch.float64,
                          device=f"cuda:{self.device}")
        check(lib().flare_sgp_predict_mean(ctx.handle, self._h, ctypes.byref(p),
                                           structure._nlist,
                                           dptr(structure.d_species), _form(form),
                                           dptr(out)))
        structure.mean_efs = out.cpu().numpy()
        return structure.mean_efs

    def predict_local_uncertainties(self, structure, form="plain"):
        """mean_efs and local variances in cluster order
        (sparse_gp.cpp:682-702; compute_cluster_uncertainties :93-133).

        The variance k_ii - |L^-1 k_i|^2 cancels, so the form matters: the
        FP64 forms ("plain", "adjoint_fp64") are exact to rounding; the FP32
        and mixed forms carry a relative variance error of up to ~7e-3 where
        std/sigma is 0.03-0.1 (55,296 atoms, 2,000 sparse), so the default
        stays FP64 for uncertainty work."""
        self._check_ready(structure)
        ctx = _capi.context(self.device)
        p = structure.descriptor_calculators[0].params()
        dev = f"cuda:{self.device}"
        out = torch.empty(1 + 3 * structure.noa + 6, dtype=torch.float64, device=dev)
        var = torch.empty(max(structure.noa, 1), dtype=torch.float64, device=dev)
        check(lib().flare_sgp_predict_local_uncertainties(
            ctx.handle, self._h, ctypes.byref(p), structure._nlist,
            dptr(structure.d_species), _form(form), dptr(out), dptr(var)))
        structure.mean_efs = out.cpu().numpy()
        structure.local_uncertainties = [var[:structure.noa].cpu().numpy()]
        return structure.mean_efs, structure.local_uncertainties

    # -- SOR / DTC (sparse_gp.cpp:630-680) --------------------------------------
    def set_R_inv(self, R_inv):
        """R^-1 of the training QR (sparse_gp.cpp:600-607); Sigma = R^-1 R^-T."""
        R = np.ascontiguousarray(np.asarray(R_inv, dtype=np.float64))
        if R.shape != (self.n_sparse, self.n_sparse):
            raise ValueError("R_inv must be n_sparse x n_sparse")
        self.R_inv = R
        self._R_handle = None  # uploaded on the next SOR / DTC prediction

    def _upload_R_inv(self):
        if getattr(self, "_R_handle", None) is self._h and self._h is not None:
            return
        ctx = _capi.context(self.device)
        t = torch.as_tensor(self.R_inv, device=f"cuda:{self.device}")
        check(lib().flare_sgp_set_R_inv(ctx.handle, self._h, dptr(t)))
        self._R_handle = self._h

    @property
    def Sigma(self):
        return self.R_inv @ self.R_inv.T

    def _struc_call(self, fn, structure, rows):
        self._build()
        ctx = _capi.context(self.device)
        p = structure.descriptor_calculators[0].params()
        n_out = 1 + 3 * structure.noa + 6
        out = torch.zeros((max(rows, 1), n_out), dtype=torch.float64,
                          device=f"cuda:{self.device}")
        check(fn(ctx.handle, self._h, ctypes.byref(p), structure._nlist,
                 dptr(structure.d_species), dptr(out)))
        return out[:rows].cpu().numpy()

    def envs_struc(self, structure):
        """Kernel matrix of the sparse set against the structure's outputs,
        n_sparse x (1+3n+6) (normalized_dot_product.cpp:115-210)."""
        return self._struc_call(lib().flare_sgp_envs_struc, structure, self.n_sparse)

    def self_kernel_struc(self, structure):
        """k(o, o) of the structure's outputs (normalized_dot_product.cpp:
        446-578)."""
        return self._struc_call(lib().flare_sgp_self_kernel_struc, structure, 1)[0]

    def _predict_var(self, fn, structure):
        self._check_ready(structure)
        if getattr(self, "R_inv", None) is None:
            raise RuntimeError("R_inv is not set")
        self._upload_R_inv()
        ctx = _capi.context(self.device)
        p = structure.descriptor_calculators[0].params()
        n_out = 1 + 3 * structure.noa + 6
        dev = f"cuda:{self.device}"
        mean = torch.empty(n_out, dtype=torch.float64, device=dev)
        var = torch.empty(n_out, dtype=torch.float64, device=dev)
        check(fn(ctx.handle, self._h, ctypes.byref(p), structure._nlist,
                 dptr(structure.d_species), dptr(mean), dptr(var)))
        structure.mean_efs = mean.cpu().numpy()
        structure.variance_efs = var.cpu().numpy()
        return structure.mean_efs, structure.variance_efs

    def predict_SOR(self, structure):
        """mean_efs and variance_efs = diag(K^T Sigma K) (sparse_gp.cpp:
        630-650)."""
        return self._predict_var(lib().flare_sgp_predict_SOR, structure)

    def predict_DTC(self, structure):
        """mean_efs and variance_efs = k(o, o) - diag(K^T K_uu^-1 K) +
        diag(K^T Sigma K) (sparse_gp.cpp:652-680)."""
        return self._predict_var(lib().flare_sgp_predict_DTC, structure)

    # -- JSON model files (sparse_gp.h:101-111) ---------------------------------
    @staticmethod
    def to_json(file_name, sgp):
        from .model_json import write_sparse_gp
        write_sparse_gp(file_name, sgp)

    @staticmethod
    def from_json(file_name, device=0):
        from .model_json import read_sparse_gp
        return read_sparse_gp(file_name, device)

    def compute_cluster_uncertainties(self, structure):
        return self.predict_local_uncertainties(structure)[1]

    # -- mapping (normalized_dot_product.cpp:643-759; sparse_gp.cpp:1142) -----
    def mapping_coefficients(self):
        self._check_ready_mapping()
        sp = self.sparse_descriptors[0]
        F, S = sp.n_descriptors, sp.n_types
        power = self.kernels[0].power
        bs = F if power == 1 else F * (F + 1) // 2
        ctx = _capi.context(self.device)
        out = torch.empty((S, bs), dtype=torch.float64, device=f"cuda:{self.device}")
        check(lib().flare_sgp_mapping_coefficients(ctx.handle, self._h, dptr(out)))
        return out.cpu().numpy()

    def _check_ready_mapping(self):
        if self.alpha is None:
            raise RuntimeError("alpha is not set")
        self._build()

    # -- variance path (normalized_dot_product.cpp:761-852;
    #    sparse_gp.cpp:1208-1396) ---------------------------------------------
    @property
    def hyperparameters(self):
        """[sigma, energy noise, force noise, stress noise]
        (sparse_gp.cpp:36-47)."""
        return np.array([self.kernels[0].sigma, self.energy_noise,
                         self.force_noise, self.stress_noise])

    @property
    def Kuu_inverse(self):
        """L^-T L^-1 (sparse_gp.cpp:588)."""
        Li = self.L_inv
        return Li.T @ Li

    def varmap_coefficients(self):
        """Variance map S x F x F of this power-1 model
        (normalized_dot_product.cpp:761-852), computed on the GPU."""
        self._build()
        sp = self.sparse_descriptors[0]
        F, S = sp.n_descriptors, sp.n_types
        ctx = _capi.context(self.device)
        out = torch.empty((S, F, F), dtype=torch.float64, device=f"cuda:{self.device}")
        check(lib().flare_sgp_varmap_coefficients(ctx.handle, self._h, dptr(out)))
        return out.cpu().numpy()

    def write_varmap_coefficients(self, file_name, contributor, kernel_index=0):
        """SparseGP::write_varmap_coefficients (sparse_gp.cpp:1208-1281);
        sets ``varmap_coeffs`` (S x F*F).  The kernel power must be 1."""
        if kernel_index != 0:
            raise ValueError("one kernel is supported")
        beta = self.varmap_coefficients()
        S = beta.shape[0]
        self.varmap_coeffs = beta.reshape(S, -1)
        calc = self.descriptor_calculators[0]
        hyps = np.ascontiguousarray(self.hyperparameters)
        cut = np.ascontiguousarray(calc.cutoffs.ravel())
        vals = np.ascontiguousarray(self.varmap_coeffs)
        pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(lib().flare_write_varmap_coefficients(
            str(file_name).encode(), str(contributor).encode(), len(hyps), pd(hyps),
            calc.n_species, calc.n_max, calc.l_max, calc.radial_basis.encode(),
            calc.cutoff_function.encode(), pd(cut), vals.shape[1], pd(vals)))

    def write_L_inverse(self, file_name, contributor):
        """SparseGP::write_L_inverse (sparse_gp.cpp:1283-1339)."""
        self._build()
        Li = np.ascontiguousarray(self.L_inv)
        calc = self.descriptor_calculators[0]
        hyps = np.ascontiguousarray(self.hyperparameters)
        cut = np.ascontiguousarray(calc.cutoffs.ravel())
        pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(lib().flare_write_L_inverse(
            str(file_name).encode(), str(contributor).encode(),
            float(self.kernels[0].power), len(hyps), pd(hyps), calc.n_species,
            calc.n_max, calc.l_max, calc.radial_basis.encode(),
            calc.cutoff_function.encode(), pd(cut), Li.shape[0], pd(Li)))

    def write_sparse_descriptors(self, file_name, contributor):
        """SparseGP::write_sparse_descriptors (sparse_gp.cpp:1341-1396)."""
        sp = self.sparse_descriptors[0]
        P, m_by_type = sp.stacked()
        P = np.ascontiguousarray(P, dtype=np.float64)
        norms = np.ascontiguousarray(np.concatenate(
            [np.asarray(n, dtype=np.float64) for n in sp.descriptor_norms]))
        arr = (ctypes.c_int * sp.n_types)(*m_by_type)
        pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(lib().flare_write_sparse_descriptors(
            str(file_name).encode(), str(contributor).encode(), sp.n_types,
            sp.n_descriptors, arr, pd(P), pd(norms)))

    def write_mapping_coefficients(self, file_name, contributor, kernel_index=0):
        from .pair import write_coefficients
        if kernel_index != 0:
            raise ValueError("one kernel is supported")
        beta = self.mapping_coefficients()
        calc = self.descriptor_calculators[0]
        write_coefficients(file_name, contributor, int(self.kernels[0].power),
                           calc.n_species, calc.n_max, calc.l_max,
                           calc.radial_basis, calc.cutoff_function, calc.cutoffs,
                           beta)
