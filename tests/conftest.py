import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ordinarydiffeq.jl_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


# GPU runs name the cached code object each compiled problem loads
os.environ.setdefault("ODEMI_LOG_CODEOBJECTS", "1")


def pytest_report_header(config):
    try:
        from odemi import _lib
        major, minor = _lib.rtc_version()
        return f"odemi: hipRTC {major}.{minor}, code-object cache {os.environ.get('ODEMI_CACHE_DIR', _lib.DEFAULT_CACHE)}"
    except Exception as e:  # the report header must not fail the session
        return f"odemi: hipRTC version unavailable ({e})"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
