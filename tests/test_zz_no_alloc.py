"""jp_count_pairs_noalloc (the entry point that allocates nothing) against
jp_count_pairs: equal counts on two calls in a row, the second on a smaller
system, so the shared scratch is reused with fewer clusters."""
import ctypes as C

import numpy as np
import pytest

from j_presto_amd import engine, native, systems

pytestmark = pytest.mark.gpu


def counts(md, fn):
    out = [C.c_int64(), C.c_int64(), C.c_int64()]
    r = getattr(md.lib, fn)(md.ctx, *[C.byref(v) for v in out])
    assert r == 0, (fn, r)
    return [v.value for v in out]


def test_noalloc_count_pairs_equals_count_pairs():
    for s in (systems.water_box(2000, 39.1, 31, n_ion_pairs=5), systems.water_box(500, 24.8, 32)):
        print(f"case noalloc n={s.n}", flush=True)
        md = engine.MdSystem(s, precision=engine.FP32)
        assert md.celprp() == 0
        a = counts(md, "jp_count_pairs")
        b = counts(md, "jp_count_pairs_noalloc")
        assert a == b and a[0] > 0 and a[2] >= a[0]
        # the state moves on: the next call counts the new coordinates the same way
        _, ier = md.runmd(12, update=10)
        assert ier == 0
        assert counts(md, "jp_count_pairs_noalloc") == counts(md, "jp_count_pairs")
        md.close()


def test_noalloc_needs_a_list():
    s = systems.water_box(200, 24.2, 33)
    md = engine.MdSystem(s)
    assert md.lib.jp_count_pairs_noalloc(md.ctx, C.byref(C.c_int64()), None, None) == native.JP_EINVAL
    md.close()
