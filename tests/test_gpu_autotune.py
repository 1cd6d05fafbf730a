"""The autotuner (vrc_autotune_stack): a bounded search on a sample of the caller's batch,
per-N winners applied to later calls, a text cache written and read back. Results under
the tuned switches equal the untuned ones to rounding (the switches choose tile shapes,
the QR kernel and the QR iteration's deflation window / bulge count, not the
mathematics)."""
import ctypes as C

import numpy as np
import pytest

from tests.test_gpu_parity_full import metric_instances
from varrcwa_amd import api
from varrcwa_amd._lib import from_dev_colmajor

pytestmark = pytest.mark.gpu


def test_autotune_sample_applies_and_caches(ctx, tmp_path):
    lib = ctx.lib
    insts = metric_instances(8, 4)
    before = from_dev_colmajor(api.batch_stack_rcwa(ctx, 5, 5, insts)["S"])
    cache = tmp_path / "tune.txt"
    arr, keep = api._stack_array(insts)
    rep = C.create_string_buffer(256)
    try:
        rc = lib.vrc_autotune_stack(ctx.h, 5, 5, len(insts), arr, 3.0, str(cache).encode(), rep, 256)
        assert rc == 0, lib.vrc_last_error(ctx.h).decode()
        text = rep.value.decode()
        print(text)
        assert text.startswith("N=50:") and "aed_window" in text and "qr_variant" in text
        assert "bulges_w48" in text and "bulges_w32" in text
        lines = [l for l in cache.read_text().splitlines() if l and not l.startswith("#")]
        assert len(lines) == 1 and lines[0].split()[0] == "50" and len(lines[0].split()) == 10
        f = lines[0].split()
        assert f[7] in ("-1", "0", "2", "4") and 1 <= int(f[8]) <= 12 and 1 <= int(f[9]) <= 7
        after = from_dev_colmajor(api.batch_stack_rcwa(ctx, 5, 5, insts)["S"])
        assert np.abs(after - before).max() <= 1e-10 * np.abs(before).max()
        # another N is not affected, and the cache reads back
        assert lib.vrc_autotune_clear() == 0
        assert lib.vrc_autotune_load(str(cache).encode()) == 1
        assert lib.vrc_autotune_load(str(tmp_path / "missing.txt").encode()) == -1
        bad = tmp_path / "bad.txt"
        bad.write_text("50 24 twelve\n")
        assert lib.vrc_autotune_load(str(bad).encode()) == -1
        bad.write_text("50 24 12 128 8 8 129 1\n")  # QR kernel 1 is not a tuning candidate
        assert lib.vrc_autotune_load(str(bad).encode()) == -1
        old = tmp_path / "seven_fields.txt"  # without the QR kernel field: the automatic choice
        old.write_text("# written before the QR kernel became a switch\n50 24 12 128 8 8 129\n")
        assert lib.vrc_autotune_load(str(old).encode()) == 1
        assert lib.vrc_autotune_stack(ctx.h, 5, 5, 0, arr, 3.0, None, None, 0) == -1
    finally:
        lib.vrc_autotune_clear()
        del keep
