"""The C-ABI library loads and exports every function include/reikna_amd.h declares
(no compute calls: runs without a GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(rk_[A-Za-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "rk_fft_plan_create" in names and "rk_fft_execute" in names
    assert len(names) > 20


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    from reikna_amd import _lib

    lib = _lib.load()
    assert hasattr(lib, name), f"{name} declared in include/reikna_amd.h but not exported"
    assert name in _lib.SIGNATURES, f"{name} has no ctypes prototype"


def test_no_device_needed_for_version():
    from reikna_amd import _lib

    assert b"gfx950" in _lib.load().rk_version()


def test_errors_without_device_are_reported():
    from reikna_amd import _lib

    lib = _lib.load()
    n = ctypes.c_int(-1)
    assert lib.rk_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0
    # a null descriptor is rejected with a message, not a crash
    assert lib.rk_fft_plan_create(None, None) == _lib.RK_ERR_INVALID
    assert "null" in _lib.last_error()


def test_missing_library_fails_loudly(tmp_path):
    """Without the HIP library every call raises: there is no CPU path."""
    import subprocess
    import sys

    code = (
        "import numpy\n"
        "from reikna_amd import ArrayMetadata\n"
        "from reikna_amd.fft import FFT\n"
        "from reikna_amd.runtime import VirtualDevice\n"
        "try:\n"
        "    FFT(ArrayMetadata((4, 64), numpy.complex64), axes=(1,)).compile(VirtualDevice())\n"
        "except Exception as e:\n"
        "    print('raised:', type(e).__name__, e)\n"
        "else:\n"
        "    print('no error')\n"
    )
    env = dict(__import__("os").environ, REIKNA_AMD_LIB=str(tmp_path / "absent.so"))
    root = __import__("pathlib").Path(__file__).resolve().parent.parent
    out = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=120)
    assert "raised: RkError" in out.stdout and "not built" in out.stdout, out.stdout + out.stderr


def test_scratch_limit_matches_header():
    """The Python constant of the no-allocation entry point's limit is the
    header's RK_SCRATCH_BYTES."""
    from reikna_amd import _lib

    text = (ROOT / "include" / "reikna_amd.h").read_text()
    m = re.search(r"#define RK_SCRATCH_BYTES \((\d+)ull << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == _lib.RK_SCRATCH_BYTES
