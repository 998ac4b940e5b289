"""flare_pp_amd -- MI355X (gfx950) force-evaluation path of flare++.

Python mirror of flare++'s ``_C_flare`` names over libflare_hip.so:
Structure, B2, DescriptorValues, ClusterDescriptor, NormalizedDotProduct,
SparseGP, plus the mapped ``pair_style flare`` potential (PairFLARE), the
``compute flare/std/atom`` uncertainty (ComputeFlareStdAtom) and
torch.autograd wrappers of the adjoint kernel.  Importing the package does
not touch the GPU; the native library is opened on first use and raises if
it is missing.
"""

__version__ = "0.1.0"

from ._capi import FlareError, FORMS  # noqa: F401
from .structure import B2, ClusterDescriptor, DescriptorValues, Structure  # noqa: F401
from .sparse_gp import NormalizedDotProduct, SparseGP  # noqa: F401
from .pair import PairFLARE, pack_beta, write_coefficients  # noqa: F401
from .calculator import (build_map, efs_stds, efs_to_results, local_stds,  # noqa: F401
                         sort_variances)
from .std_atom import ComputeFlareStdAtom, read_std_files  # noqa: F401
