"""The seeded sweep generator of batched runs (CPU only)."""
from mpifdtd_amd import batch as B


def test_splitmix64_known_values():
    # SplitMix64 outputs for seed 0 (Steele, Lea and Flood's reference sequence)
    g = B.SplitMix64(0)
    assert [g.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_sweep_is_reproducible_and_in_range():
    a, b, c = B.sweep_spec(7, 64), B.sweep_spec(7, 64), B.sweep_spec(8, 64)
    assert a == b and a != c
    for s in a:
        assert s["side"] in B.SIDES and s["width_nm"] == (s["side"] + 32) * 5
        assert 380 <= s["lambda_nm"] <= 780 and 6000 <= s["steps"] <= 10000
        p = s["params"]
        if s["kind"] == "multilayer":
            assert 1.0 <= p["ml_n0"] ** 2 <= 4.0 and 1.0 <= p["ml_n1"] ** 2 <= 4.0
            assert 50 <= p["ml_thick_nm0"] <= 300 and 1 <= p["ml_layer_num"] <= 20
            assert (p["ml_thick_nm0"] + p["ml_thick_nm1"]) * p["ml_layer_num"] <= 0.8 * s["side"] * 5
        else:
            assert 1.5 <= p["zz_n"] ** 2 <= 3.0 and 200 <= p["zz_width_nm"] <= 800 and 50 <= p["zz_thick_nm"] <= 300
    kinds = {s["kind"] for s in a}
    assert kinds == {"multilayer", "zigzag"}
