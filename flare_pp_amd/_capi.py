"""ctypes binding of libflare_hip.so (include/flare_hip.h).

torch is imported before the library is opened so that the library binds to
the HIP runtime torch already loaded (same sonames); device buffers handed
across are torch tensors' data pointers.
"""

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime first; see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflare_hip.so")

FLARE_OK = 0
ERRORS = {1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "EIO", 5: "EFORMAT"}

MAX_SPECIES = 8
RADIAL_CHEBYSHEV = 0
CUTOFF_QUADRATIC = 0
CUTOFF_COSINE = 1
CUTOFF_CODES = {"quadratic": CUTOFF_QUADRATIC, "cosine": CUTOFF_COSINE}

FORM_PLAIN_FP64 = 0
FORM_ADJOINT_FP64 = 1
FORM_ADJOINT_FP32 = 2
FORM_ADJOINT_MIXED = 3
FORMS = {
    "plain": FORM_PLAIN_FP64,
    "plain_fp64": FORM_PLAIN_FP64,
    "adjoint_fp64": FORM_ADJOINT_FP64,
    "adjoint_fp32": FORM_ADJOINT_FP32,
    "adjoint_mixed": FORM_ADJOINT_MIXED,
}


class FlareError(RuntimeError):
    """A libflare_hip call returned an error status."""

    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class B2Params(ctypes.Structure):
    _fields_ = [
        ("n_species", ctypes.c_int),
        ("n_max", ctypes.c_int),
        ("l_max", ctypes.c_int),
        ("radial_basis", ctypes.c_int),
        ("cutoff_function", ctypes.c_int),
        ("radial_r1", ctypes.c_double),
        ("cutoffs", ctypes.POINTER(ctypes.c_double)),
    ]


class NListView(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int),
        ("n_pairs", ctypes.c_int64),
        ("sweep", ctypes.c_int),
        ("volume", ctypes.c_double),
        ("single_sweep_cutoff", ctypes.c_double),
        ("cutoff", ctypes.c_double),
        ("species_count", ctypes.c_int * MAX_SPECIES),
        ("d_neighbor_count", ctypes.c_void_p),
        ("d_cumulative_neighbor_count", ctypes.c_void_p),
        ("d_structure_indices", ctypes.c_void_p),
        ("d_neighbor_species", ctypes.c_void_p),
        ("d_relative_positions", ctypes.c_void_p),
        ("d_wrapped_positions", ctypes.c_void_p),
        ("d_reverse_pair", ctypes.c_void_p),
        ("d_atom_indices", ctypes.c_void_p),
    ]


class PairInfo(ctypes.Structure):
    _fields_ = [
        ("power", ctypes.c_int),
        ("n_species", ctypes.c_int),
        ("n_max", ctypes.c_int),
        ("l_max", ctypes.c_int),
        ("n_descriptors", ctypes.c_int),
        ("beta_size", ctypes.c_int64),
        ("cutoff_function", ctypes.c_int),
        ("legacy", ctypes.c_int),
        ("cutoff", ctypes.c_double),
        ("radial_basis", ctypes.c_char * 32),
        ("cutoff_name", ctypes.c_char * 32),
    ]


class StdInfo(ctypes.Structure):
    _fields_ = [
        ("use_map", ctypes.c_int),
        ("power", ctypes.c_int),
        ("n_hyps", ctypes.c_int),
        ("hyps", ctypes.c_double * 4),
        ("n_species", ctypes.c_int),
        ("n_max", ctypes.c_int),
        ("l_max", ctypes.c_int),
        ("n_descriptors", ctypes.c_int),
        ("beta_size", ctypes.c_int64),
        ("n_kernels", ctypes.c_int),
        ("n_clusters", ctypes.c_int),
        ("n_clusters_by_type", ctypes.c_int * MAX_SPECIES),
        ("cutoff_function", ctypes.c_int),
        ("cutoff", ctypes.c_double),
        ("radial_basis", ctypes.c_char * 32),
        ("cutoff_name", ctypes.c_char * 32),
    ]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_cp = ctypes.c_char_p
_pvp = ctypes.POINTER(ctypes.c_void_p)
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every entry point of include/flare_hip.h
SIGNATURES = {
    "flare_version": (_cp, []),
    "flare_last_error": (_cp, []),
    "flare_ctx_create": (_i, [_i, _pvp]),
    "flare_ctx_destroy": (_i, [_vp]),
    "flare_ctx_set_stream": (_i, [_vp, _vp]),
    "flare_ctx_synchronize": (_i, [_vp]),
    "flare_ctx_set_blocking": (_i, [_vp, _i]),
    "flare_ctx_profile": (_i, [_vp, _i]),
    "flare_ctx_profile_get": (_i, [_vp, _i, ctypes.c_char_p, _i, _pd, _pi64]),
    "flare_nlist_create": (_i, [_vp, _pvp]),
    "flare_nlist_destroy": (_i, [_vp]),
    "flare_nlist_build": (_i, [_vp, _vp, _i, _pd, _vp, _vp, _d]),
    "flare_nlist_get_view": (_i, [_vp, ctypes.POINTER(NListView)]),
    "flare_b2_n_descriptors": (_i, [ctypes.POINTER(B2Params)]),
    "flare_b2_compute": (_i, [_vp, ctypes.POINTER(B2Params), _vp, _vp, _vp, _vp,
                              _vp]),
    "flare_b2_count_pairs": (_i, [_vp, ctypes.POINTER(B2Params), _vp, _vp, _pi64,
                                  _pi64]),
    "flare_b2_compute_dervs": (_i, [_vp, ctypes.POINTER(B2Params), _vp, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _vp]),
    "flare_pair_load": (_i, [_vp, _cp, _pvp]),
    "flare_coeff_read": (_i, [_cp, ctypes.POINTER(PairInfo), _pd, _pd]),
    "flare_pair_create": (_i, [_vp, _i, _i, _i, _i, _i, _pd, _pd, _pvp]),
    "flare_pair_destroy": (_i, [_vp]),
    "flare_pair_get_info": (_i, [_vp, ctypes.POINTER(PairInfo)]),
    "flare_pair_get_cutoffs": (_i, [_vp, _pd]),
    "flare_pair_get_beta": (_i, [_vp, _vp, _pd]),
    "flare_pair_compute": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "flare_pair_compute_noalloc": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp,
                                        _vp]),
    "flare_pair_compute_neigh": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp, _i64, _i64,
                                      _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "flare_pair_beta_grad": (_i, [_vp, _vp, _vp]),
    "flare_write_coefficients": (_i, [_cp, _cp, _i, _i, _i, _i, _cp, _cp, _pd,
                                      _i64, _pd]),
    "flare_sgp_create": (_i, [_vp, _i, _i, _pi, _vp, _d, _d, _d, _pvp]),
    "flare_sgp_destroy": (_i, [_vp]),
    "flare_sgp_set_alpha": (_i, [_vp, _vp, _vp]),
    "flare_sgp_get_L_inverse": (_i, [_vp, _vp, _vp]),
    "flare_sgp_get_Kuu": (_i, [_vp, _vp, _vp]),
    "flare_sgp_predict_mean": (_i, [_vp, _vp, ctypes.POINTER(B2Params), _vp, _vp,
                                    _i, _vp]),
    "flare_sgp_predict_local_uncertainties": (_i, [_vp, _vp,
                                                   ctypes.POINTER(B2Params), _vp,
                                                   _vp, _i, _vp, _vp]),
    "flare_envs_envs": (_i, [_vp, _i, _i, _pi, _vp, _pi, _vp, _d, _d, _vp]),
    "flare_sgp_mapping_coefficients": (_i, [_vp, _vp, _vp]),
    "flare_synth_rsa": (_i, [_i, _d, _d, ctypes.c_uint64, ctypes.c_int64, _vp]),
    "flare_ctx_set_plain_chunk": (_i, [_vp, _i64]),
    "flare_std_read": (_i, [_cp, _cp, ctypes.POINTER(StdInfo), _pd, _pd, _pd]),
    "flare_std_load": (_i, [_vp, _cp, _cp, _pvp]),
    "flare_std_create": (_i, [_vp, ctypes.POINTER(StdInfo), _pd, _pd, _pd, _pvp]),
    "flare_std_destroy": (_i, [_vp]),
    "flare_std_get_info": (_i, [_vp, ctypes.POINTER(StdInfo)]),
    "flare_std_compute": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "flare_sgp_varmap_coefficients": (_i, [_vp, _vp, _vp]),
    "flare_sgp_envs_struc": (_i, [_vp, _vp, ctypes.POINTER(B2Params), _vp, _vp, _vp]),
    "flare_sgp_self_kernel_struc": (_i, [_vp, _vp, ctypes.POINTER(B2Params), _vp, _vp,
                                         _vp]),
    "flare_sgp_set_R_inv": (_i, [_vp, _vp, _vp]),
    "flare_sgp_predict_SOR": (_i, [_vp, _vp, ctypes.POINTER(B2Params), _vp, _vp, _vp,
                                   _vp]),
    "flare_sgp_predict_DTC": (_i, [_vp, _vp, ctypes.POINTER(B2Params), _vp, _vp, _vp,
                                   _vp]),
    "flare_write_varmap_coefficients": (_i, [_cp, _cp, _i, _pd, _i, _i, _i, _cp,
                                             _cp, _pd, _i64, _pd]),
    "flare_write_L_inverse": (_i, [_cp, _cp, _d, _i, _pd, _i, _i, _i, _cp, _cp,
                                   _pd, _i, _pd]),
    "flare_write_sparse_descriptors": (_i, [_cp, _cp, _i, _i, _pi, _pd, _pd]),
}

_lib = None


def lib():
    """Open libflare_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`make -C flare_pp_amd/csrc` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != FLARE_OK:
        raise FlareError(rc, lib().flare_last_error().decode())


def dptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def dbl_array(values):
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr


class Context:
    """Owns a flare_ctx bound to one device and torch's current stream."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("libflare_hip needs a ROCm GPU (none visible)")
        self.device = device
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            check(lib().flare_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.sync_stream()

    def sync_stream(self):
        """Enqueue on torch's current stream of this device."""
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().flare_ctx_set_stream(self.handle, ctypes.c_void_p(s)))

    def set_blocking(self, blocking):
        check(lib().flare_ctx_set_blocking(self.handle, int(bool(blocking))))

    def set_plain_chunk(self, nbytes):
        """Budget of the plain form's per-pair derivative tensor (bytes)."""
        check(lib().flare_ctx_set_plain_chunk(self.handle, int(nbytes)))

    def synchronize(self):
        check(lib().flare_ctx_synchronize(self.handle))

    def profile(self, enable):
        check(lib().flare_ctx_profile(self.handle, int(bool(enable))))

    def stage_times(self):
        """{stage: (total_ms, count)} recorded since profiling was enabled."""
        out = {}
        k = 0
        buf = ctypes.create_string_buffer(64)
        tot = ctypes.c_double()
        cnt = ctypes.c_int64()
        while lib().flare_ctx_profile_get(self.handle, k, buf, 64, ctypes.byref(tot),
                                          ctypes.byref(cnt)) == FLARE_OK:
            out[buf.value.decode()] = (tot.value, cnt.value)
            k += 1
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            _lib.flare_ctx_destroy(h)
            self.handle = None


_contexts = {}


def context(device=0):
    """Process-wide context of a device."""
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    ctx.sync_stream()
    return ctx
