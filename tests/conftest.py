import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libruqu_gpu.so on the device)")


def _ensure_built():
    """Build the oracle and the product library in-tree when they are missing
    (needs gcc and hipcc; a machine without them uses the prebuilt files)."""
    import subprocess
    oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(oracle_so):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib_so = os.path.join(ROOT, "ruqu_amd", "libruqu_gpu.so")
    if not os.path.exists(lib_so) and os.path.exists("/opt/rocm/bin/hipcc"):
        subprocess.run([sys.executable, os.path.join(ROOT, "ruqu_amd", "build.py")], check=True)


_ensure_built()
