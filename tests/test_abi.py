"""The C-ABI library: it loads, exports every symbol include/pht_splat.h declares, and the ctypes
mirror has the C struct layouts (checked against a probe compiled from the header). CPU only:
no compute entry point is called here."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pht_splat.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pht_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from photara_studio_amd import _abi
    assert _abi.LIB_PATH.exists(), "build the kernels first (__graft_entry__.build())"
    lib = C.CDLL(str(_abi.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds exactly the declared entry points
    assert sorted(_abi.SIGNATURES) == names


STRUCTS = ["pht_camera", "pht_settings", "pht_gaussians", "pht_view_outputs", "pht_view_grads", "pht_model_grads",
           "pht_adam", "pht_capacity", "pht_counts", "pht_loss_config", "pht_densify_stats",
           "pht_densify_observation", "pht_row_buffer", "pht_split_params", "pht_densify_step"]


def test_struct_layouts_match_header(tmp_path):
    from photara_studio_amd import _abi
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pht_splat.h"', "int main(void){"]
    for s in STRUCTS:
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f, _ in getattr(_abi, s)._fields_:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c_vals = dict(line.split() for line in out if line.strip())
    for s in STRUCTS:
        cls = getattr(_abi, s)
        assert int(c_vals[s]) == C.sizeof(cls), s
        for f, _ in cls._fields_:
            assert int(c_vals[f"{s}.{f}"]) == getattr(cls, f).offset, f"{s}.{f}"


def test_workspace_bytes_host_only():
    from photara_studio_amd import _abi
    lib = _abi.load()
    assert lib.pht_abi_version() == 4
    cap = _abi.pht_capacity(10000, 1, 16, 256, 65536, 1 << 20)
    b1 = lib.pht_workspace_bytes(C.byref(cap))
    cap2 = _abi.pht_capacity(10000, 1, 16, 256, 65536, 1 << 21)
    b2 = lib.pht_workspace_bytes(C.byref(cap2))
    assert b1 > 0 and b2 - b1 >= 16 * (1 << 20)  # four u32 instance arrays grow with the capacity
    bad = _abi.pht_capacity(10000, 1, 12, 256, 65536, 1 << 20)  # tile size 12 is not supported
    assert lib.pht_workspace_bytes(C.byref(bad)) == 0
    assert b"tile_size" in lib.pht_last_error()
    # invalid workspace creation is rejected before touching a device
    h = C.c_void_p()
    assert lib.pht_workspace_create(None, 0, C.byref(cap), C.byref(h)) == _abi.PHT_EINVAL


def test_select_host_helpers():
    """Host-only parts of the row selection: workspace size and the value behind a key."""
    import struct
    from photara_studio_amd import _abi
    lib = _abi.load()
    assert lib.pht_select_workspace_bytes(1000) >= 8 * 1000 + 1024
    for v in (0.25, -3.5, 0.0, 2.0 ** -140):
        bits = struct.unpack("<I", struct.pack("<f", v))[0]
        order = bits ^ (0xFFFFFFFF if bits & 0x80000000 else 0x80000000)
        assert lib.pht_select_key_value((order << 32) | 5, _abi.PHT_SELECT_LARGEST) == pytest.approx(v, abs=0)
        assert lib.pht_select_key_value(((~order & 0xFFFFFFFF) << 32) | 5, _abi.PHT_SELECT_SMALLEST) == \
            pytest.approx(v, abs=0)
    # bad arguments are rejected before any launch
    ws = C.create_string_buffer(64)
    assert lib.pht_select_rows(None, None, None, 10, 1, 7, 0, None, 1, None, None, ws, 64, None) == _abi.PHT_EINVAL
    assert lib.pht_select_rows(None, None, None, 10, 1, 0, 0, None, 1, None, None, ws, 64, None) == _abi.PHT_EINVAL
    assert b"workspace" in lib.pht_last_error()


def test_product_path_fails_loudly_without_library(monkeypatch, tmp_path):
    from photara_studio_amd import _abi
    monkeypatch.setattr(_abi, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(_abi, "_lib", None)
    with pytest.raises(ImportError):
        _abi.load()
