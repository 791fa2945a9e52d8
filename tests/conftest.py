"""Test configuration: import paths and the `gpu` marker.

`-m "not gpu"` runs on any host (oracle vs golden vectors, host logic, the
C-ABI export table); `-m gpu` needs an MI355X and calls the HIP path through
the C ABI.
"""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "pursuit-evasion-core_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available() -> bool:
    try:
        import pe_amd._native as n

        n.device_info()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    if not oracle.LIB_PATH.exists():
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    import pe_amd._native as n

    info = n.device_info()  # raises loudly without a gfx950 device
    print(f"[gpu] {info['name']} CUs={info['compute_units']}", flush=True)
    return info
