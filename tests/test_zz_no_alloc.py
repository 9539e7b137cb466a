"""cgo_inventory_noalloc (the allocation-free inventory) through the C ABI:
the same sums as cgo_inventory, bit for bit, on two calls in a row (the
second on a member with fewer tracers, so the static scratch array holds the
first call's sums beyond its length), and a clear refusal above the
scratch array's size."""
import numpy as np
import pytest

from cgenie_amd import engine as E
from cases import make_case

pytestmark = pytest.mark.gpu


def _member(eng, c, prec):
    gid = eng.add_grid(c.g, c.k1)
    m = eng.add_member(gid, c.tset.L, prec, c.diff)
    eng.set_velocity(m, c.u)
    eng.upload(m, c.ts)
    return m


def test_noalloc_inventory_equals_inventory():
    eng = E.Engine(0)
    try:
        big = make_case(kind="config2", seed=71)                  # 24 tracers
        small = make_case(18, 18, 8, kind="ts", L=5, seed=72)     # 5 tracers
        a = _member(eng, big, E.FP64)
        b = _member(eng, small, E.FP32C)
        eng.step([a, b], 3)
        for m in (a, b):
            want = eng.inventory(m)
            got = eng.inventory_noalloc(m)
            print("member", m, "tracers", len(want), "identical:", np.array_equal(got, want), flush=True)
            assert got.shape == want.shape and np.array_equal(got, want)
        # a change to the member table (a new member) since the last step
        c = _member(eng, small, E.FP64)
        assert np.array_equal(eng.inventory_noalloc(c), eng.inventory(c))
        wide = make_case(18, 18, 8, kind="ts", L=65, seed=73)
        w = _member(eng, wide, E.FP64)
        with pytest.raises(E.CgoError, match="at most 64"):
            eng.inventory_noalloc(w)
    finally:
        eng.close()
