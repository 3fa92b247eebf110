import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
for p in (ROOT, ORACLE_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs with -m gpu on the GPU box)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (oracle/pyoracle.py); builds oracle/_build if it is missing."""
    lib = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
    cpu = os.path.join(ORACLE_DIR, "_build", "hpcscan_cpu")
    if not (os.path.exists(lib) and os.path.exists(cpu)):
        subprocess.run(["make", "-C", ORACLE_DIR, "-j8"], check=True, capture_output=True)
    import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def capi():
    """The MI355X C ABI (hpcscan_amd.capi); fails loudly if the library is not built."""
    from hpcscan_amd import capi as c

    c.lib()
    return c


def run_bin(path, args, cwd, timeout=600):
    """Run a harness binary; return (returncode, stdout)."""
    r = subprocess.run([path] + list(args), cwd=cwd, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def report_value(out: str, label: str):
    """Value printed by printInfo for `label` (first match), as a string."""
    for line in out.splitlines():
        if line.strip().startswith(label):
            return line.strip()[len(label):].strip()
    return None
