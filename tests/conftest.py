import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libptl_hip.so)")


def _ensure_cpu_libs():
    """Builds the host library and the oracle when they are missing (CPU only;
    the HIP library is built by __graft_entry__.build())."""
    from ptilopsis_amd import build
    if not os.path.exists(os.path.join(ROOT, "ptilopsis_amd", "libptl_host.so")):
        build.build_host()
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build.build_oracle()


_ensure_cpu_libs()


@pytest.fixture(scope="session")
def gpu_ctx():
    from ptilopsis_amd.generator import Context
    ctx = Context()
    yield ctx
    ctx.close()
