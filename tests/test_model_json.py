"""SparseGP JSON model files (sparse_gp.h:101-111, sparse_gp.cpp:1406-1417).
No JSON file written by the reference is available (parity unpinned): the
CPU test reads a file laid out from the nlohmann field lists; the GPU test
checks that a written model predicts exactly as the original."""
import json

import numpy as np
import pytest

from flare_pp_amd import SparseGP


def test_read_hand_laid_file(tmp_path):
    F = 1 * 2 * 3 // 2 * 2  # B2 {1, 2, 1}: nr = 2, F = 3 * 2 = 6
    rows = [[0.1 * (i + 1) + 0.01 * k for k in range(F)] for i in range(3)]
    j = {
        "hyperparameters": [1.5, 0.01, 0.02, 0.003],
        "kernels": [{"kernel_name": "NormalizedDotProduct", "sigma": 1.5, "sig2": 2.25,
                     "power": 2.0, "kernel_hyperparameters": [1.5]}],
        "Kuu_jitter": 1e-8, "alpha": [0.5, -1.0, 2.0],
        "R_inv": [[1.0, 0.1, 0.0], [0.0, 1.0, 0.2], [0.0, 0.0, 0.5]],
        "energy_noise": 0.01, "force_noise": 0.02, "stress_noise": 0.003,
        "sparse_descriptors": [{"descriptors": [rows], "descriptor_norms":
                                [[float(np.linalg.norm(r)) for r in rows]],
                                "cutoff_values": [[1.0, 1.0, 1.0]],
                                "n_clusters_by_type": [3], "cumulative_type_count": [0],
                                "n_descriptors": F, "n_types": 1, "n_clusters": 3}],
        "training_structures": [{"descriptor_calculators": [{
            "descriptor_name": "B2", "radial_basis": "chebyshev",
            "cutoff_function": "quadratic", "radial_hyps": [0.0, 4.5],
            "cutoff_hyps": [], "descriptor_settings": [1, 2, 1],
            "cutoffs": [[4.5]]}]}],
        "sparse_indices": [[[0, 4, 7]]],
    }
    p = tmp_path / "model.json"
    p.write_text(json.dumps(j))
    gp = SparseGP.from_json(p)
    assert gp.kernels[0].sigma == 1.5 and gp.kernels[0].power == 2.0
    assert (gp.energy_noise, gp.force_noise, gp.stress_noise) == (0.01, 0.02, 0.003)
    assert gp.n_sparse == 3
    np.testing.assert_array_equal(gp.alpha, [0.5, -1.0, 2.0])
    np.testing.assert_array_equal(gp.R_inv, np.array(j["R_inv"]))
    P, m = gp.sparse_descriptors[0].stacked()
    np.testing.assert_array_equal(P, np.array(rows))
    b2 = gp.descriptor_calculators[0]
    assert b2.descriptor_settings == [1, 2, 1] and b2.cutoffs[0, 0] == 4.5
    np.testing.assert_array_equal(gp.hyperparameters, j["hyperparameters"])


@pytest.mark.gpu
def test_written_model_predicts_the_same(cfg1, tmp_path):
    from flare_pp_amd import B2, NormalizedDotProduct, Structure
    c = cfg1
    b2 = B2("chebyshev", "quadratic", [0, 5.0], [], c["settings"])
    s = Structure(c["cell"], c["species"], c["positions"], 5.0, [b2])
    gp = SparseGP([NormalizedDotProduct(c["sigma"], 2)], 0.01, 0.02, 0.003)
    gp.add_specific_environments(s, c["sparse_atoms"][:20])
    gp.set_alpha(c["alpha"][:20])
    rng = np.random.default_rng(0)
    gp.set_R_inv(np.triu(rng.normal(0, 0.05, (20, 20))) + np.eye(20))
    p = tmp_path / "gp.json"
    SparseGP.to_json(p, gp)
    gp2 = SparseGP.from_json(p)
    a = gp.predict_local_uncertainties(s)
    b = gp2.predict_local_uncertainties(s)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1][0], b[1][0])
    np.testing.assert_array_equal(gp.predict_DTC(s)[1], gp2.predict_DTC(s)[1])
    np.testing.assert_array_equal(gp.mapping_coefficients(), gp2.mapping_coefficients())
    np.testing.assert_array_equal(gp.L_inv, gp2.L_inv)
