"""Validation and performance summaries in the reference's tab-separated
formats (Validation/Validation.jl:7-80, 358-407, 588-646), so results of
this build can be diffed with the reference's own tooling.

* `ValidationResult(name, description, metrics, tolerances)`: passed when
  every |metric| is within its tolerance (Validation.jl:76-80).
* `write_validation_summary` / `read_validation_summary` /
  `compare_validation_summaries` (Validation.jl:358-407): header
  `case passed metric value tolerance description`, one row per metric in
  sorted metric order.
* `performance_metric` / `write_performance_summary` /
  `read_performance_summary` (Validation.jl:588-646): header
  `case operation seconds bytes description`.

Numbers are written as Julia prints a Float64 (`1.0e-14`, `0.001`,
`100000.0`, `Inf`), booleans as `true`/`false`.
"""
import math
import time
from dataclasses import dataclass, field


def julia_float(x):
    """A Float64 as Julia's `print` writes it: the shortest round-trip
    digits, fixed notation for decimal exponents -5 < e < 6 (with `.0` for
    integral values), `d.ddde±x` otherwise, `Inf`, `-Inf`, `NaN`."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    digits, exp = _shortest(x)
    sign = "-" if x < 0 else ""
    e10 = exp + len(digits) - 1  # decimal exponent of the leading digit
    if -5 < e10 < 6:
        if e10 >= 0:
            whole = digits[:e10 + 1].ljust(e10 + 1, "0")
            frac = digits[e10 + 1:] or "0"
        else:
            whole, frac = "0", "0" * (-e10 - 1) + digits
        return f"{sign}{whole}.{frac}"
    mant = digits[0] + "." + (digits[1:] or "0")
    return f"{sign}{mant}e{e10}"


def _shortest(x):
    """(digit string, exponent) with x = ±0.digits... shifted: value = int(digits) * 10**exp,
    using Python's shortest round-trip repr."""
    r = repr(abs(x))
    if "e" in r or "E" in r:
        mant, e = r.lower().split("e")
        e = int(e)
    else:
        mant, e = r, 0
    if "." in mant:
        a, b = mant.split(".")
    else:
        a, b = mant, ""
    b = b.rstrip("0")
    digits = (a + b).lstrip("0")
    exp = e - len(b)
    stripped = digits.rstrip("0")
    exp += len(digits) - len(stripped)
    return stripped or "0", exp


def _parse_float(s):
    return float(s)


@dataclass
class ValidationResult:
    name: str
    description: str
    metrics: dict
    tolerances: dict
    passed: bool = field(init=False)

    def __post_init__(self):
        self.metrics = {str(k): float(v) for k, v in self.metrics.items()}
        self.tolerances = {str(k): float(v) for k, v in self.tolerances.items()}
        self.passed = all(abs(v) <= self.tolerances[k] for k, v in self.metrics.items())


def validation_passed(results):
    if isinstance(results, ValidationResult):
        return results.passed
    return all(getattr(r, "passed") for r in results)


def write_validation_summary(path, results):
    with open(path, "w") as io:
        io.write("case\tpassed\tmetric\tvalue\ttolerance\tdescription\n")
        for r in results:
            for metric in sorted(r.metrics):
                io.write(f"{r.name}\t{'true' if r.passed else 'false'}\t{metric}\t{julia_float(r.metrics[metric])}\t"
                         f"{julia_float(r.tolerances[metric])}\t{r.description}\n")
    return path


def read_validation_summary(path):
    """{(case, metric): value} (Validation.jl:375-390); ValueError on a bad file."""
    with open(path) as io:
        lines = io.read().splitlines()
    if not lines:
        raise ValueError(f"validation summary is empty: {path}")
    if lines[0].split("\t") != ["case", "passed", "metric", "value", "tolerance", "description"]:
        raise ValueError(f"validation summary has an unexpected header: {lines[0]}")
    data = {}
    for line in lines[1:]:
        if not line.strip():
            continue
        parts = line.split("\t")
        if len(parts) != 6:
            raise ValueError(f"validation summary row must have 6 columns: {line}")
        data[(parts[0], parts[2])] = _parse_float(parts[3])
    return data


@dataclass
class ExternalComparisonResult:
    case: str
    metric: str
    reference: float
    candidate: float
    absolute_error: float
    relative_error: float
    passed: bool


def compare_validation_summaries(reference_path, candidate_path, atol=0.0, rtol=1e-12):
    """Validation.jl:392-407: every reference metric against the candidate's
    (NaN when missing)."""
    ref = read_validation_summary(reference_path)
    cand = read_validation_summary(candidate_path)
    out = []
    eps = 2.220446049250313e-16
    for key in sorted(ref, key=lambda k: str(k)):
        r = ref[key]
        c = cand.get(key, math.nan)
        ae = abs(c - r)
        re = ae / max(abs(r), eps)
        passed = math.isfinite(c) and ae <= atol + rtol * abs(r)
        out.append(ExternalComparisonResult(key[0], key[1], r, c, ae, re, passed))
    return out


@dataclass
class PerformanceMetric:
    case: str
    operation: str
    seconds: float
    bytes: float
    description: str


def performance_metric(func, case, operation, description, nbytes=0.0):
    """Validation.jl:588-592: one untimed call, then one timed call.  Julia
    reports the bytes its call allocated; here `nbytes` (default 0) is
    recorded, e.g. the HBM bytes the operation moves."""
    func()
    t0 = time.perf_counter()
    func()
    return PerformanceMetric(str(case), str(operation), time.perf_counter() - t0, float(nbytes), description)


def write_performance_summary(path, metrics):
    with open(path, "w") as io:
        io.write("case\toperation\tseconds\tbytes\tdescription\n")
        for m in metrics:
            io.write(f"{m.case}\t{m.operation}\t{julia_float(m.seconds)}\t{julia_float(m.bytes)}\t{m.description}\n")
    return path


def read_performance_summary(path):
    with open(path) as io:
        lines = io.read().splitlines()
    if not lines:
        raise ValueError(f"performance summary is empty: {path}")
    if lines[0].split("\t") != ["case", "operation", "seconds", "bytes", "description"]:
        raise ValueError(f"performance summary has an unexpected header: {lines[0]}")
    out = []
    for line in lines[1:]:
        if not line.strip():
            continue
        parts = line.split("\t")
        if len(parts) != 5:
            raise ValueError(f"performance summary row must have 5 columns: {line}")
        out.append(PerformanceMetric(parts[0], parts[1], _parse_float(parts[2]), _parse_float(parts[3]), parts[4]))
    return out


# --- validation cases run on the device --------------------------------------

@dataclass
class ValidationCase:
    name: str
    description: str
    run: object


def constant_action_validation():
    """Validation.jl:115-128 on a polar grid: no transport and no sources,
    uniform action: zero tendency and no change over an RK3 step."""
    import numpy as np
    from .grids import RectilinearGrid, PolarWaveVectorGrid
    from .model import SpectralWaveModel, set_, time_step_
    g = RectilinearGrid(size=(3, 3, 2), x=(0, 3), y=(0, 3), z=(-1, 0))
    sg = PolarWaveVectorGrid(kappa=[0.4, 0.8], phi=[0.0, 2.1, 4.2])
    m = SpectralWaveModel(g, sg, horizontal_advection=None, timestepper="RK3")
    set_(m, N=1.0)
    initial = m.action.copy()
    G = m.tendency()
    time_step_(m, 0.25)
    metrics = {"max_tendency": float(np.abs(G).max()), "state_change": float(np.abs(m.action - initial).max())}
    return ValidationResult("constant_action", "source-only no-op action", metrics,
                            {"max_tendency": 0.0, "state_change": 0.0})


def q_normalization_validation():
    """Validation.jl:150-176 through the device Doppler contraction: a
    depth-uniform unit current projects to U = sum_k W = 1 for every kappa,
    and dU/dkappa = 0."""
    import numpy as np
    from .grids import RectilinearGrid, PolarWaveVectorGrid
    from .model import SpectralWaveModel
    g = RectilinearGrid(size=(3, 3, 8), x=(0, 3), y=(0, 3), z=(-1, 0))
    sg = PolarWaveVectorGrid(kappa=[1e-8, 0.25, 1.0, 4.0], phi=[0.0, 2.1, 4.2])
    ones = np.ones((3, 3, 8))
    m = SpectralWaveModel(g, sg, velocities=(ones, ones), timestepper="RK3")
    Ux, _ = m.doppler_velocity()
    dUx, _ = m.doppler_velocity_kappa_derivative()
    metrics = {"normalization_error": float(np.abs(Ux - 1).max()),
               "derivative_normalization_error": float(np.abs(dUx).max())}
    return ValidationResult("q_normalization", "Q kernel vertical normalization", metrics,
                            {"normalization_error": 1e-12, "derivative_normalization_error": 1e-10})


def intrinsic_conservation_validation():
    """Fused intrinsic WENO5 transport on a periodic grid telescopes: the
    total action (compensated FP64 sum) is unchanged by RK3 steps while no
    cell goes negative."""
    import numpy as np
    from .diagnostics import total_action
    from .grids import RectilinearGrid, PolarWaveVectorGrid
    from .model import SpectralWaveModel, set_, time_step_
    g = RectilinearGrid(size=(32, 24, 1), x=(0, 32.0), y=(0, 24.0), z=(-1, 0))
    sg = PolarWaveVectorGrid(kappa=np.geomspace(0.2, 1.5, 6), phi=np.linspace(0, 2 * math.pi, 13)[:12])
    m = SpectralWaveModel(g, sg, timestepper="RK3")
    rng = np.random.default_rng(0)
    set_(m, N=rng.uniform(0.1, 1.0, size=m.size))
    before = total_action(m)
    time_step_(m, 0.05, 10)
    after = total_action(m)
    metrics = {"relative_action_change": abs(after - before) / before,
               "negative_cells": float((m.action < 0).sum())}
    return ValidationResult("intrinsic_conservation", "periodic intrinsic transport conserves action", metrics,
                            {"relative_action_change": 1e-12, "negative_cells": 0.0})


def default_validation_cases():
    return (ValidationCase("constant_action", "A source-only model with uniform action does not change.",
                           constant_action_validation),
            ValidationCase("q_normalization", "The Q kernel integrates to one over the water column.",
                           q_normalization_validation),
            ValidationCase("intrinsic_conservation", "Fused intrinsic transport conserves total action.",
                           intrinsic_conservation_validation))


def run_validation(cases=None):
    cases = default_validation_cases() if cases is None else cases
    return [c.run() for c in cases]
