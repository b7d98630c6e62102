"""The C-ABI library loads and exports every symbol include/hermeshd_amd.h
declares (CPU: no compute calls), and the host-side config logic."""
import ctypes
import os
import re

import pytest

from hermeshd_amd import _lib
from hermeshd_amd.config import HDConfig, sod_1d, vortex_3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hd_\w+)\s*\(", txt, re.M))
    return names


def test_header_declares_expected_entry_points():
    names = declared_symbols()
    assert set(_lib.EXPORTED) <= names, set(_lib.EXPORTED) - names
    assert names <= set(_lib.EXPORTED) | {n for n in names if n.startswith("hd_ens_") or n.startswith("hd_adj")}


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_config_struct_layout(tmp_path):
    """The ctypes mirror of hd_config matches the C layout (offsetof from gcc)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    fields = [f[0] for f in _lib.HDConfigC._fields_]
    src = tmp_path / "probe.c"
    body = "\n".join(f'printf("{f} %zu\\n", offsetof(hd_config, {f}));' for f in fields)
    src.write_text(f'#include <stdio.h>\n#include <stddef.h>\n#include "hermeshd_amd.h"\n'
                   f'int main(void){{ {body} printf("sizeof %zu\\n", sizeof(hd_config)); return 0; }}\n')
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(line.split() for line in subprocess.check_output([str(exe)]).decode().splitlines())
    for f in fields:
        assert int(out[f]) == getattr(_lib.HDConfigC, f).offset, f
    assert int(out["sizeof"]) == ctypes.sizeof(_lib.HDConfigC)


def test_config_validation():
    with pytest.raises(ValueError):
        HDConfig(nx=4, ny=4, nz=4, nbasis=10, iquad=2)
    with pytest.raises(ValueError):
        HDConfig(nx=4, ny=4, nz=4, bc=("periodic", "wall", "periodic", "periodic", "periodic", "periodic"))
    with pytest.raises(ValueError):
        HDConfig(nx=4, ny=4, nz=4, ivis=1, nQ=5)
    with pytest.raises(ValueError):
        HDConfig(nx=4, ny=4, nz=4, ivis=2, nQ=5)
    with pytest.raises(ValueError):
        HDConfig(nx=4, ny=4, nz=4, ivis=2, nQ=11, noise=True)
    with pytest.raises(ValueError):
        HDConfig(nx=4, ny=4, nz=4, ivis=3, nQ=11)
    HDConfig(nx=4, ny=4, nz=4, ivis=2, nQ=11)
    c = vortex_3d()
    assert c.shape == (10, 5, 32, 128, 128) and c.cflm == 0.08
    assert abs(1.0 / c.dxi - 10.0 / 128) < 1e-15
    assert sod_1d().cflm == 0.3


def test_min_dt_static_refuses_large_grids_before_any_device_call():
    """The size check of the allocation-free CFL entry point runs on the host,
    before any HIP call (so it runs here without a GPU): 256x256x257 cells
    need more workgroup maxima than its static scratch holds."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    cs = _lib.config_struct(HDConfig(nx=256, ny=256, nz=257, nbasis=4, iquad=2, nQ=5).as_ints())
    dummy = ctypes.c_void_p(8)  # never dereferenced: the call returns before touching it
    assert L.hd_min_dt_static(ctypes.byref(cs), dummy, dummy, None) == _lib.HD_ERR_INVALID
    assert b"16777216" in L.hd_last_error()
    assert L.hd_min_dt_static(None, dummy, dummy, None) == _lib.HD_ERR_INVALID


def test_create_refuses_13_moment_grids_past_32_bit_face_indices():
    """The 13-moment kernels index cells and faces in 32 bits, so hd_create
    refuses a grid with (nx+1)(ny+1)(nz+1), a bound on the faces of every
    direction, above 2^31 - 1; the check runs before any HIP call (so it runs
    here without a GPU).  1290^3 cells give 1291^3 = 2.152e9."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    from hermeshd_amd.config import taylor_green_3d
    L = _lib.lib()
    cfg = taylor_green_3d(1290)
    assert 1291 * 1290 * 1290 > 2 ** 31 - 1
    cs = _lib.config_struct(cfg.as_ints())
    out = ctypes.c_void_p()
    assert L.hd_create(ctypes.byref(cs), ctypes.byref(out)) == _lib.HD_ERR_INVALID
    assert b"2^31" in L.hd_last_error() and not out.value
