import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "dopamine-fdm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

ORACLE_LIB = os.path.join(ROOT, "oracle", "lib", "libdfp_oracle.so")
DFP_LIB = os.path.join(ROOT, "dopamine-fdm_amd", "lib", "libdfp.so")


def _build_if_missing(path, makedir):
    # builds only when nothing is built; after a source change run
    # __graft_entry__.build() (which always rebuilds)
    if not os.path.exists(path):
        subprocess.run(["make", "-j", "8", "-C", makedir], check=True, stdout=subprocess.DEVNULL)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: large CPU oracle runs")


@pytest.fixture(scope="session", autouse=True)
def built_libraries():
    _build_if_missing(ORACLE_LIB, os.path.join(ROOT, "oracle"))
    _build_if_missing(DFP_LIB, os.path.join(ROOT, "dopamine-fdm_amd", "csrc"))
