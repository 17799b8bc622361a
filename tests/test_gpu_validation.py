"""The device validation cases (constant action, Q normalization, intrinsic
conservation) pass and write a reference-format summary."""
import pytest

from ripple_amd import validation as V

pytestmark = pytest.mark.gpu


def test_validation_suite(tmp_path):
    print("[gpu-test] validation suite", flush=True)
    results = V.run_validation()
    assert [r.name for r in results] == ["constant_action", "q_normalization", "intrinsic_conservation"]
    for r in results:
        print(f"[gpu-test] {r.name}: {r.metrics}", flush=True)
    assert V.validation_passed(results)
    path = tmp_path / "validation_summary.tsv"
    V.write_validation_summary(path, results)
    parsed = V.read_validation_summary(path)
    assert parsed[("constant_action", "max_tendency")] == 0.0
    assert V.validation_passed(V.compare_validation_summaries(path, path))
