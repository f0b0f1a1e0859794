"""The seeded uneven microstructure family (host-side parameters only)."""


def test_family_parameters_are_deterministic_and_in_range():
    from jaxmaterials_amd.data.batch import GRIDS, sample_params

    ps = [sample_params(i) for i in range(384)]
    assert ps[17] == sample_params(17)
    assert {p["grid"] for p in ps} == set(GRIDS)
    for p in ps:
        assert 2 <= p["layers"] <= 8 and 1.5 <= p["radius"] <= 5.0 and 0.3 <= p["vf"] <= 0.6
        assert 5.0 <= p["mu_ratio"] <= 100.0
        assert abs(p["lam_matrix"] - 2.5927) < 1e-3
