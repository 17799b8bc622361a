"""The validation and performance summary formats of
Validation/Validation.jl:358-407, 594-646 (CPU only): headers, row shapes,
Julia's number printing, round trips, comparison and error handling, as
test/validation/validation_suite.jl checks them."""
import math

import pytest

from ripple_amd import validation as V


def _results():
    return [V.ValidationResult("constant_action", "source-only no-op action",
                               {"max_tendency": 0.0, "state_change": 0.0}, {"max_tendency": 0.0, "state_change": 0.0}),
            V.ValidationResult("q_normalization", "Q kernel vertical normalization",
                               {"normalization_error": 2.220446049250313e-16, "derivative_normalization_error": 1e-17},
                               {"normalization_error": 1e-12, "derivative_normalization_error": 1e-10}),
            V.ValidationResult("failing", "a metric above its tolerance", {"err": 0.5}, {"err": 0.1})]


def test_julia_number_printing():
    cases = {1e-14: "1.0e-14", 0.001: "0.001", 1e-4: "0.0001", 1e-5: "1.0e-5", 1e5: "100000.0",
             123456.0: "123456.0", 1e6: "1.0e6", 1234567.0: "1.234567e6", -2.25e-8: "-2.25e-8", 0.0: "0.0",
             math.inf: "Inf", 3.0: "3.0", 1e16: "1.0e16", 0.1: "0.1", 2.220446049250313e-16: "2.220446049250313e-16"}
    for x, s in cases.items():
        assert V.julia_float(x) == s, (x, V.julia_float(x), s)
        assert float(s) == x


def test_validation_summary_round_trip(tmp_path):
    results = _results()
    assert [r.passed for r in results] == [True, True, False]
    assert not V.validation_passed(results) and V.validation_passed(results[:2])
    path = tmp_path / "validation_summary.tsv"
    V.write_validation_summary(path, results)
    lines = path.read_text().splitlines()
    assert lines[0] == "case\tpassed\tmetric\tvalue\ttolerance\tdescription"
    assert len(lines) == 1 + sum(len(r.metrics) for r in results)
    assert all(len(line.split("\t")) == 6 for line in lines)
    assert lines[1].split("\t")[:2] == ["constant_action", "true"]
    parsed = V.read_validation_summary(path)
    assert parsed[("constant_action", "max_tendency")] == 0.0
    assert parsed[("q_normalization", "normalization_error")] == 2.220446049250313e-16
    comparison = V.compare_validation_summaries(path, path)
    assert V.validation_passed(comparison) and all(c.absolute_error == 0 for c in comparison)
    perturbed = tmp_path / "perturbed.tsv"
    perturbed.write_text(path.read_text().replace("0.0", "1.0", 1))
    assert not V.validation_passed(V.compare_validation_summaries(path, perturbed, atol=0.0, rtol=0.0))


def test_summary_errors(tmp_path):
    empty = tmp_path / "empty.tsv"
    empty.write_text("")
    with pytest.raises(ValueError):
        V.read_validation_summary(empty)
    bad = tmp_path / "bad.tsv"
    bad.write_text("case\tvalue\n")
    with pytest.raises(ValueError):
        V.read_validation_summary(bad)
    short = tmp_path / "short.tsv"
    short.write_text("case\tpassed\tmetric\tvalue\ttolerance\tdescription\nx\ttrue\tm\n")
    with pytest.raises(ValueError):
        V.read_validation_summary(short)
    with pytest.raises(ValueError):
        V.read_performance_summary(bad)


def test_performance_summary_round_trip(tmp_path):
    calls = []
    m = V.performance_metric(lambda: calls.append(1), "product_field", "set_and_m0", "Set and sum.", nbytes=1.5e9)
    assert len(calls) == 2 and m.seconds >= 0 and m.bytes == 1.5e9
    path = tmp_path / "performance.tsv"
    V.write_performance_summary(path, [m, V.PerformanceMetric("sources", "step", 0.25, 0.0, "A step.")])
    lines = path.read_text().splitlines()
    assert lines[0] == "case\toperation\tseconds\tbytes\tdescription"
    assert lines[2] == "sources\tstep\t0.25\t0.0\tA step."
    parsed = V.read_performance_summary(path)
    assert [p.case for p in parsed] == ["product_field", "sources"]
    assert parsed[0].bytes == 1.5e9
