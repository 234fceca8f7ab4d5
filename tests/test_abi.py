"""The C-ABI library loads and exports every entry point include/chs.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "chs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(chs_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "chs_chebyshev_filter" in names and "chs_rayleigh_ritz" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "chemrust-scf_amd", "libchs_gpu.so"))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_is_not_linked_into_the_product():
    so = open(os.path.join(ROOT, "chemrust-scf_amd", "libchs_gpu.so"), "rb").read()
    assert b"liboracle" not in so and b"orc_" not in so


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus above 1 exits with status 2 and says so,
    before anything touches a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
