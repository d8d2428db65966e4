"""The C-ABI library loads and exports every symbol include/vapor_mac3d.h
declares; struct layouts agree; defaults and presets are the reference's. No
compute calls (runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

import oracle
from vapor_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "vapor_mac3d.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vgx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(abi.LIBRARY_PATH))
    names = header_symbols()
    assert len(names) >= 35
    missing = [name for name in names if not hasattr(lib, name)]
    assert not missing, missing
    bound = {name for name, _, _ in abi.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_library_exports_every_simplicial_symbol():
    """include/vapor_simplicial3d.h: every declaration is exported."""
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "vapor_simplicial3d.h").read_text(), flags=re.S)
    names = sorted(set(re.findall(r"\b(vsx_[a-z0-9_]+)\s*\(", text)))
    assert len(names) >= 11
    lib = ctypes.CDLL(str(abi.LIBRARY_PATH))
    assert not [name for name in names if not hasattr(lib, name)]


def test_library_exports_every_mac2d_symbol():
    """include/vapor_mac2d.h: every declaration exported and bound, struct
    sizes agree, defaults equal the reference's (where it is built)."""
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "vapor_mac2d.h").read_text(), flags=re.S)
    names = sorted(set(re.findall(r"\b(vgx2r?_[a-z0-9_]+)\s*\(", text)))
    assert len(names) >= 40
    lib = ctypes.CDLL(str(abi.LIBRARY_PATH))
    assert not [name for name in names if not hasattr(lib, name)]
    assert set(names) == {name for name, _, _ in abi.SIGNATURES_2D}
    lib = abi.load_library()
    for index, struct in enumerate(abi.STRUCTS_2D):
        assert lib.vgx2_struct_size(index) == ctypes.sizeof(struct)
    from oracle import mac2d as O2
    if O2.ref2d_available():
        for fn, ref in ((lib.vgx2_params_default, O2.ref2d_lib().r2_params_default),
                        (lib.vgx2_params_plume, O2.ref2d_lib().r2_params_plume)):
            a, b = abi.Vgx2Params(), abi.Vgx2Params()
            fn(ctypes.byref(a))
            assert ref(ctypes.byref(b)) == 0
            assert as_bytes(a) == as_bytes(b)
        for n, method in ((16, -1), (128, 0), (128, 1)):
            a, b = abi.Vgx2rParams(), abi.Vgx2rParams()
            if method < 0:
                lib.vgx2r_params_default(ctypes.byref(a))
            else:
                assert lib.vgx2r_params_plume(method, ctypes.byref(a)) == 0
            assert O2.ref2d_lib().r2h_params(n, method, ctypes.byref(b)) == 0
            assert as_bytes(a) == as_bytes(b)


def test_struct_layouts_agree():
    lib = abi.load_library()
    for index, struct in enumerate(abi.STRUCTS):
        assert lib.vgx_struct_size(index) == ctypes.sizeof(struct)
    assert ctypes.sizeof(abi.VgxParams) == 136
    assert ctypes.sizeof(abi.VgxProjectionOptions) == 64
    assert ctypes.sizeof(abi.VgxProjectionResult) == 80


def as_bytes(s):
    return bytes(memoryview(s))


def test_defaults_equal_oracle_and_reference():
    lib = abi.load_library()
    p = abi.VgxParams()
    lib.vgx_params_default(ctypes.byref(p))
    assert as_bytes(p) == as_bytes(oracle.default_params())
    o = abi.VgxProjectionOptions()
    lib.vgx_projection_options_default(ctypes.byref(o))
    assert as_bytes(o) == as_bytes(oracle.default_options())
    if oracle.ref_available():
        r = abi.VgxParams()
        oracle.ref_lib().ref_params_default(ctypes.byref(r))
        for name, _ in abi.VgxParams._fields_[:19]:
            assert getattr(p, name) == getattr(r, name), name


def test_presets_match_reference():
    lib = abi.load_library()
    for algorithm in range(4):
        p = abi.VgxParams()
        assert lib.vgx_params_preset(algorithm, ctypes.byref(p)) == abi.VGX_OK
        # MacGridFluidSolver3D.cpp:11-26
        assert p.reflection == (algorithm >= 2)
        assert p.maccormack_smoke == (algorithm % 2 == 1) == p.maccormack_vel
        assert (p.emitter_radius, p.source_strength, p.buoyancy) == (pytest.approx(0.08), 4.0, 3.0)
        assert p.project_iterations == 600 and p.sor_omega == pytest.approx(1.9)
        if oracle.ref_available():
            r = abi.VgxParams()
            assert oracle.ref_lib().ref_params_preset(algorithm, ctypes.byref(r)) == 0
            for name, _ in abi.VgxParams._fields_[:19]:
                assert getattr(p, name) == getattr(r, name), (algorithm, name)
    p = abi.VgxParams()
    assert lib.vgx_params_preset(7, ctypes.byref(p)) == abi.VGX_EINVAL
    assert b"Unknown MAC algorithm" in lib.vgx_last_error()


def test_product_never_uses_the_oracle():
    for path in (ROOT / "vapor_amd").rglob("*"):
        if path.suffix in (".py", ".hip", ".hpp", ".cpp", ".h"):
            text = path.read_text()
            assert "import oracle" not in text and "from oracle" not in text, path
            assert "mac3d_oracle" not in text and "orc_" not in text, path


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        abi.load_library(tmp_path / "absent.so")
