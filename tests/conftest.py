import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


# Under MOJOBLAS_SANITIZED_BUILD=1 (set by tests/test_sanitized_cpu.py for
# its child run) the C-ABI library and the oracle are the host-sanitized
# builds under _build/san instead of the shipped ones.
if os.environ.get("MOJOBLAS_SANITIZED_BUILD") == "1":
    import oracle as _oracle
    from mojoblas_amd import _lib as _mlib

    _san = os.path.join(ROOT, "_build", "san")
    _mlib.LIB_PATH = os.path.join(_san, "libmojoblas_hip.so")
    _oracle._LIB_PATH = os.path.join(_san, "liboracle.so")


def pytest_sessionfinish(session, exitstatus):
    if os.environ.get("MOJOBLAS_SANITIZED_BUILD") == "1":
        with open("/proc/self/maps") as f:
            maps = f.read()
        san = os.path.join(ROOT, "_build", "san")
        if all(os.path.join(san, lib) in maps for lib in ("libmojoblas_hip.so", "liboracle.so")):
            print("\nSANITIZED_LIBS_LOADED")
