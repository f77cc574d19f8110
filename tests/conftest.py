import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "permutationalinvariantdynamics.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpimi on device 0)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    import parity_log
    if not parity_log.RECORDS:
        return
    terminalreporter.write_line("measured parity errors (value / bar):")
    for name, value, bar in parity_log.RECORDS:
        terminalreporter.write_line(f"  {name} {value:.2e} / {bar:.0e}")
