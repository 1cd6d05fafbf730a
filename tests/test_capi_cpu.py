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


def test_find_guided_modes_host_entry_point():
    """vrc_find_guided_modes runs on the host (NULL context): the oracle's selection and order
    on random spectra with repeated real parts, and rejection of k0 <= 0."""
    import numpy as np

    from oracle import rcwa_oracle as O
    from varrcwa_amd import api
    rng = np.random.default_rng(4)
    k0 = 2 * O.PI / 1.55
    g = k0 * k0 * (rng.uniform(0.5, 4.0, 200) + 1j * rng.uniform(-0.5, 0.5, 200))
    g[10] = g[20] = g[30]  # ties keep index order
    assert api.find_guided_modes(None, g, k0, 0.2, 1.445, 3.48) == O.find_guided_modes(g, k0, 0.2, 1.445, 3.48)
    assert api.find_guided_modes(None, g[:0], k0, 0.2, 1.445, 3.48) == []
    lib = _lib.load()
    cnt = C.c_int(0)
    assert lib.vrc_find_guided_modes(None, 1, g.ctypes.data, -1.0, 0.2, 1.0, 2.0, None, None, C.byref(cnt)) == -1


def test_gds_reader_host_entry_point(tmp_path):
    """vrc_gds_polygons runs the library's GDSII reader on the host: the same polygons, bit for
    bit, as the oracle's reader on the reference's layout file and on a written file (closing
    point, BOX, PATH and second cell), and an error for a missing file."""
    import os

    import pytest

    from oracle import rcwa_oracle as O
    from tests.gds_files import synthetic_cells, write_gds
    from varrcwa_amd import api
    f = tmp_path / "t.gds"
    write_gds(f, synthetic_cells())
    fixture = os.path.join(os.path.dirname(__file__), "golden", "contour_extract.gds")
    for path in (str(f), fixture):
        ref = O.read_gds_polygons(path)
        for cell in range(len(ref)):
            assert api.gds_polygons(path, 1e-6, cell) == ref[cell]
    with pytest.raises(api.VrcError):
        api.gds_polygons(str(tmp_path / "missing.gds"))


def test_tuning_switches_and_cache_on_the_host(tmp_path):
    """vrc_tune_set validates keys and values, and the autotuner's cache file is parsed on the
    host: seven-field lines (automatic QR kernel), eight-field lines and ten-field lines (with the
    bulge counts of QR kernels 2 and 4) load; a QR kernel that is not a tuning candidate, a bulge
    count out of range, a ninth field without a tenth, trailing text, a malformed line or a
    missing file is an error. No device call."""
    lib = _lib.load()
    try:
        assert lib.vrc_tune_set(b"no.such.switch", 1) == -1
        assert lib.vrc_tune_set(None, 1) == -1
        assert lib.vrc_tune_set(b"zgeev.qr_variant", 3) == -1
        assert lib.vrc_tune_set(b"zgeev.qr_variant", 5) == -1
        assert lib.vrc_tune_set(b"zgeev.qr_variant", 4) == 0
        assert lib.vrc_tune_set(b"zgeev.qr_variant", -1) == 0
        assert lib.vrc_tune_set(b"zgeev.qr_variant", -2) == -1
        assert lib.vrc_tune_set(b"zgeev.qr_variant", 2) == 0
        assert lib.vrc_tune_set(b"zgeev.qr_variant", 0) == 0
        assert lib.vrc_tune_set(b"zgeev.bulges_w48", 8) == 0
        good = tmp_path / "tune.txt"
        good.write_text("# a comment\n882 24 12 128 8 8 129\n242 16 10 64 8 4 65 2\n98 16 6 64 8 4 65 4\n"
                        "450 24 12 128 8 8 129 2 10 5\n")
        assert lib.vrc_autotune_load(str(good).encode()) == 4
        for bad in ("882 24 12 128 8 8\n", "882 24 12 128 8 8 129 1\n", "1 24 12 128 8 8 129\n", "x y\n",
                    "882 24 12 128 8 8 129 2 8\n", "882 24 12 128 8 8 129 2 13 6\n",
                    "882 24 12 128 8 8 129 2 8 0\n", "882 24 12 128 8 8 129 2 8 6 extra\n",
                    "882 24 12 128 8 8 129 two\n"):
            f = tmp_path / "bad.txt"
            f.write_text(bad)
            assert lib.vrc_autotune_load(str(f).encode()) == -1, bad
        assert lib.vrc_autotune_load(str(tmp_path / "missing.txt").encode()) == -1
        assert lib.vrc_autotune_load(None) == -1
    finally:
        assert lib.vrc_autotune_clear() == 0
        lib.vrc_tune_set(b"zgeev.qr_variant", -1)
