"""Result conventions of flare++'s ASE calculator (sparse_gp_calculator.py)
without the ASE dependency: the Python caller contract of the GP path."""

import numpy as np


def efs_to_results(mean_efs, n_atoms, single_atom_energies=None, species=None):
    """energy, forces (n x 3) and ASE-order stress from mean_efs
    (sparse_gp_calculator.py:56-77): ASE stress = -[s0, s3, s5, s4, s2, s1]
    of flare's (xx, xy, xz, yy, yz, zz)."""
    efs = np.asarray(mean_efs)
    energy = float(efs[0])
    if single_atom_energies is not None:
        for s in species:
            energy += single_atom_energies[s]
    forces = efs[1:1 + 3 * n_atoms].reshape(-1, 3).copy()
    s = efs[1 + 3 * n_atoms:]
    stress = -np.array([s[0], s[3], s[5], s[4], s[2], s[1]])
    return energy, forces, stress


def sort_variances(atom_indices, variances):
    """Cluster-order variances back to atom order (sparse_gp_calculator.py:
    166-184); atom_indices is DescriptorValues.atom_indices."""
    n = len(variances)
    out = np.zeros(n)
    v = 0
    for block in atom_indices:
        for a in block:
            out[a] = variances[v]
            v += 1
    return out


def local_stds(atom_indices, variances, sigma):
    """Signed square roots of the atom-order variances divided by |sigma|
    (sparse_gp_calculator.py:94-109); shape (n, 3), column 0 filled."""
    sv = sort_variances(atom_indices, variances)
    stds = np.where(sv > 0, np.sqrt(np.abs(sv)), -np.sqrt(np.abs(sv)))
    full = np.zeros((len(sv), 3))
    full[:, 0] = stds / abs(sigma)
    return full


def efs_stds(variance_efs):
    """Per-force-component stds of the SOR / DTC variances (n x 3): signed
    square roots of variance_efs[1:-6] (sparse_gp_calculator.py:81-94)."""
    v = np.asarray(variance_efs)[1:-6]
    return np.where(v > 0, np.sqrt(np.abs(v)), -np.sqrt(np.abs(v))).reshape(-1, 3)


def build_map(sparse_gp, filename="lmp.flare", contributor="user"):
    """SGP_Calculator.build_map (sparse_gp_calculator.py:159-163): the
    pair_style coefficient file plus the L^-1 and sparse-descriptor files of
    compute flare/std/atom (L_inv_<filename>, sparse_desc_<filename>)."""
    import os
    d, b = os.path.split(str(filename))
    sparse_gp.write_mapping_coefficients(filename, contributor, 0)
    sparse_gp.write_L_inverse(os.path.join(d, f"L_inv_{b}"), contributor)
    sparse_gp.write_sparse_descriptors(os.path.join(d, f"sparse_desc_{b}"),
                                       contributor)
