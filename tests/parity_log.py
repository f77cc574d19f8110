"""Measured parity errors of the GPU tests, printed in one block at the end of
the session (tests/conftest.py), so a passing run records the numbers and
not only pass/fail."""

RECORDS = []


def record(name, value, bar):
    """Keep `value` (a measured relative error) beside the bar the test asserts."""
    RECORDS.append((name, float(value), float(bar)))
    print(f"parity {name}: {float(value):.3e} (bar {float(bar):.0e})", flush=True)
    return value
