"""CPU-side checks of the C ABI library (no GPU calls): it loads and exports every
entry point that include/varrcwa.h declares, and argument validation rejects bad input
before touching the device."""
import ctypes as C

from varrcwa_amd import _lib


def test_library_exports_every_declared_symbol():
    have, declared = _lib.exported_symbols()
    missing = sorted(set(declared) - set(have))
    assert not missing, missing
    assert len(declared) >= 30


def test_version_string():
    lib = _lib.load()
    assert b"gfx950" in lib.vrc_version()


def test_null_context_is_rejected():
    lib = _lib.load()
    assert lib.vrc_zgemm(None, 0, 0, 1, 1, 1, None, None, 1, 1, None, 1, 1, None, None, 1, 1, 1) == -1
    assert lib.vrc_sync(None) == -1
    assert lib.vrc_integrator_destroy(None) == 0
    assert lib.vrc_sampler_ndim(None) == -1
    out = C.c_void_p()
    assert lib.vrc_ctx_create(0, None, None) == -1
    assert out.value is None
