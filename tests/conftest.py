import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    from pigfem_amd import _ffi
    if not os.path.exists(_ffi.LIB_PATH):
        from pigfem_amd.build import build
        build()
    from oracle import oracle
    if not os.path.exists(oracle.SO):
        oracle.build()
