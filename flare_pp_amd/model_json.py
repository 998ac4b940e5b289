"""SparseGP JSON model files: the nlohmann serialisation of SparseGP
(sparse_gp.h:101-108, sparse_gp.cpp:1406-1417) read as a model for
prediction, and written with the same keys.

Layout (nlohmann field names; Eigen vectors are arrays, matrices arrays of
rows, an empty matrix is null; json.h:8-80):
  hyperparameters   [sigma, energy noise, force noise, stress noise]
  kernels           [{kernel_name, sigma, sig2, power,
                      kernel_hyperparameters}]  (normalized_dot_product.h:69)
  sparse_descriptors[{descriptors (per type: rows), descriptor_norms,
                      cutoff_values, n_clusters_by_type,
                      cumulative_type_count, n_descriptors, n_types,
                      n_clusters}]  (descriptor.h:87-89)
  training_structures[{..., descriptor_calculators: [{descriptor_name "B2",
                      radial_basis, cutoff_function, radial_hyps,
                      cutoff_hyps, descriptor_settings, cutoffs}]}]
                      (structure.h:104-110, b2.cpp:548-571)
  alpha, L_inv, R_inv, Sigma, Kuu, Kuu_inverse, Kuu_jitter, noises, counts.
Reading takes the kernel, the sparse set, the B2 settings of the first
training structure, alpha and R_inv; L^-1 is recomputed on the device from
the sparse set (the same K_uu + jitter factorisation).  Training data and
likelihood terms are written empty: training is not part of this library.
"""

import json

import numpy as np

from .structure import B2, ClusterDescriptor


def _vec(x):
    return [] if x is None else [float(v) for v in np.asarray(x).ravel()]


def _mat(x):
    if x is None:
        return None
    a = np.asarray(x, dtype=np.float64)
    if a.size == 0:
        return None
    return [[float(v) for v in row] for row in a]


def _arr(x, shape=None):
    if x is None:
        return np.zeros(shape if shape else (0,))
    a = np.asarray(x, dtype=np.float64)
    return a.reshape(shape) if shape is not None and a.size == np.prod(shape) else a


def _b2_json(b2):
    return {"descriptor_name": "B2", "radial_basis": b2.radial_basis,
            "cutoff_function": b2.cutoff_function,
            "radial_hyps": list(b2.radial_hyps), "cutoff_hyps": list(b2.cutoff_hyps),
            "descriptor_settings": list(b2.descriptor_settings),
            "cutoffs": _mat(b2.cutoffs)}


def _structure_stub(calcs):
    # keys of Structure (structure.h:104-110); only the descriptor
    # calculators carry information here
    keys_mat = ["cell", "cell_transpose", "cell_transpose_inverse", "cell_dot",
                "cell_dot_inverse", "positions", "wrapped_positions",
                "relative_positions"]
    d = {k: None for k in keys_mat}
    d.update({"neighbor_count": [], "cumulative_neighbor_count": [],
              "structure_indices": [], "neighbor_species": [], "cutoff": 0.0,
              "single_sweep_cutoff": 0.0, "volume": 0.0, "sweep": 0,
              "n_neighbors": 0, "species": [], "noa": 0, "energy": [],
              "forces": [], "stresses": [], "mean_efs": [], "variance_efs": [],
              "mean_contributions": [], "local_uncertainties": [],
              "descriptor_calculators": [_b2_json(c) for c in calcs],
              "descriptors": []})
    return d


def write_sparse_gp(file_name, sgp):
    """SparseGP::to_json (sparse_gp.cpp:1406-1410) with the keys of
    sparse_gp.h:101-108."""
    k = sgp.kernels[0]
    sp = sgp.sparse_descriptors[0]
    L_inv = sgp.L_inv if sgp.n_sparse else None
    Kuu = sgp.Kuu if sgp.n_sparse else None
    R = getattr(sgp, "R_inv", None)
    out = {
        "hyperparameters": _vec(sgp.hyperparameters),
        "kernels": [{"kernel_name": "NormalizedDotProduct", "sigma": k.sigma,
                     "sig2": k.sig2, "power": k.power,
                     "kernel_hyperparameters": _vec(k.kernel_hyperparameters)}],
        "Kuu_kernels": [_mat(Kuu)], "Kuf_kernels": [], "Kuu": _mat(Kuu), "Kuf": None,
        "n_kernels": 1, "Kuu_jitter": sgp.Kuu_jitter,
        "Sigma": _mat(R @ R.T) if R is not None else None,
        "Kuu_inverse": _mat(L_inv.T @ L_inv) if L_inv is not None else None,
        "R_inv": _mat(R), "L_inv": _mat(L_inv),
        "alpha": _vec(sgp.alpha),
        "R_inv_diag": _vec(np.diag(R)) if R is not None else [],
        "L_diag": _vec(np.diag(L_inv)) if L_inv is not None else [],
        "sparse_descriptors": [{
            "descriptors": [[[float(v) for v in row] for row in np.asarray(d).reshape(
                -1, sp.n_descriptors)] for d in sp.descriptors],
            "descriptor_norms": [_vec(n) for n in sp.descriptor_norms],
            "cutoff_values": [_vec(c) for c in sp.cutoff_values],
            "n_clusters_by_type": list(map(int, sp.n_clusters_by_type)),
            "cumulative_type_count": list(map(int, sp.cumulative_type_count)),
            "n_descriptors": int(sp.n_descriptors), "n_types": int(sp.n_types),
            "n_clusters": int(sp.n_clusters)}],
        "training_structures": [_structure_stub(sgp.descriptor_calculators)],
        "sparse_indices": [[list(map(int, a)) for a in sgp.sparse_indices[0]]],
        "noise_vector": [], "y": [], "label_count": [], "n_energy_labels": 0,
        "n_force_labels": 0, "n_stress_labels": 0, "n_sparse": int(sgp.n_sparse),
        "n_labels": 0, "n_strucs": 0, "energy_noise": sgp.energy_noise,
        "force_noise": sgp.force_noise, "stress_noise": sgp.stress_noise,
        "log_marginal_likelihood": 0.0, "data_fit": 0.0, "complexity_penalty": 0.0,
        "trace_term": 0.0, "constant_term": 0.0, "likelihood_gradient": [],
    }
    with open(file_name, "w") as f:
        json.dump(out, f)


def read_sparse_gp(file_name, device=0):
    """SparseGP::from_json (sparse_gp.cpp:1412-1417) as a prediction model."""
    from .sparse_gp import NormalizedDotProduct, SparseGP
    with open(file_name) as f:
        j = json.load(f)
    kj = j["kernels"]
    if len(kj) != 1 or kj[0].get("kernel_name") != "NormalizedDotProduct":
        raise ValueError("one NormalizedDotProduct kernel is supported")
    hyps = j.get("hyperparameters") or []
    kern = NormalizedDotProduct(kj[0]["sigma"], kj[0]["power"])
    gp = SparseGP([kern], j.get("energy_noise", hyps[1] if len(hyps) > 1 else 0.0),
                  j.get("force_noise", hyps[2] if len(hyps) > 2 else 0.0),
                  j.get("stress_noise", hyps[3] if len(hyps) > 3 else 0.0),
                  device=device)
    gp.Kuu_jitter = float(j.get("Kuu_jitter", gp.Kuu_jitter))
    calcs = []
    for st in j.get("training_structures") or []:
        for dc in st.get("descriptor_calculators") or []:
            if dc.get("descriptor_name") == "B2":
                S = int(dc["descriptor_settings"][0])
                cut = np.asarray(dc["cutoffs"], dtype=np.float64).reshape(S, S) \
                    if dc.get("cutoffs") else None
                calcs.append(B2(dc["radial_basis"], dc["cutoff_function"],
                                dc["radial_hyps"], dc["cutoff_hyps"],
                                dc["descriptor_settings"], cut))
        if calcs:
            break
    gp.descriptor_calculators = calcs[:1]
    spj = j["sparse_descriptors"][0]
    c = ClusterDescriptor()
    c.initialize_cluster(int(spj["n_types"]), int(spj["n_descriptors"]))
    F = c.n_descriptors
    c.descriptors = [np.asarray(d, dtype=np.float64).reshape(-1, F) if d else
                     np.zeros((0, F)) for d in spj["descriptors"]]
    c.descriptor_norms = [np.asarray(v, dtype=np.float64) for v in spj["descriptor_norms"]]
    c.cutoff_values = [np.asarray(v, dtype=np.float64) for v in spj["cutoff_values"]]
    c.n_clusters_by_type = list(map(int, spj["n_clusters_by_type"]))
    c.cumulative_type_count = list(map(int, spj["cumulative_type_count"]))
    c.n_clusters = int(spj["n_clusters"])
    gp.sparse_descriptors = [c]
    si = j.get("sparse_indices") or [[]]
    gp.sparse_indices = [list(si[0])]
    if j.get("alpha"):
        gp.set_alpha(np.asarray(j["alpha"], dtype=np.float64))
    if j.get("R_inv"):
        gp.set_R_inv(np.asarray(j["R_inv"], dtype=np.float64))
    return gp
