"""The rounding-sensitivity variants of the oracle (oracle/sensitivity.py):
each builds, the default build reproduces itself exactly, and every variant
changes the FP32 solve (so each switch reaches the arithmetic it names) while
staying within FP32 noise of the FP64 twin on a small case."""


def test_variants_change_only_rounding():
    from oracle import sensitivity
    rows = sensitivity.main(["--size", "12", "--schedule", "2,2,4", "--variants", "dot_abc,fma,weight_ulp"])
    by = {r["variant"]: r for r in rows}
    assert by["fp32"]["vs_fp32_oracle"]["max_abs"] == 0.0
    for name in ("dot_abc", "fma", "weight_ulp"):
        d = by[name]["vs_fp32_oracle"]["max_abs"]
        assert 0.0 < d < 1e-4, (name, d)
        assert by[name]["vs_fp64"]["max_abs"] < 1e-4
