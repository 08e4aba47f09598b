"""The C-ABI library loads on the CPU host and exports every entry point the
header declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from vork_amd import abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vork_hip.h")).read()
    return sorted(set(re.findall(r"\b(vh_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_abi_version_and_defaults():
    L = native.lib()
    assert L.vh_abi_version() == abi.VH_ABI_VERSION == 2
    p = abi.RenderParams()
    L.vh_render_params_default(ctypes.byref(p), 1920, 1080, 64)
    assert (p.width, p.height, p.samples, p.sample_end) == (1920, 1080, 64, 64)
    assert (p.max_bounces, p.diffuse_bounces, p.glossy_bounces, p.caustics) == (12, 4, 4, 3)
    assert abs(p.clamp_indirect - 10.0) < 1e-6 and abs(p.light_threshold - 0.01) < 1e-9
    assert p.clamp_direct == 0.0


def test_usage_errors_need_no_device():
    L = native.lib()
    out = ctypes.c_void_p()
    assert L.vh_device_create(0, None) == abi.VH_ERR_USAGE
    buf = ctypes.create_string_buffer(256)
    assert L.vh_last_error(buf, 256) == abi.VH_ERR_USAGE
    assert b"null" in buf.value
    assert L.vh_render(None, None, 0, None, None) == abi.VH_ERR_USAGE
    del out


def test_rust_binding_in_the_integration_notes_covers_every_entry_point():
    """The Rust binding written out for the reference side binds exactly the
    header's entry points (one `pub fn` per symbol, none missing or extra)."""
    import os
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    text = open(os.path.join(here, "..", "INTEGRATION.md")).read()
    bound = set(re.findall(r"pub fn (vh_\w+)\(", text))
    assert bound == set(declared_symbols())
