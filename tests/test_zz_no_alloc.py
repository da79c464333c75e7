"""The allocation-free entry point fp4_np_4ph_modes_noalloc against
fp4_np_4ph_modes, through the C ABI.

Bar: counts and phase-space sums bit for bit equal (both run the same count
and reduction kernels in the same order), on two calls in a row, the second
with a smaller mode subset than the first (so stale scratch contents from the
first call would show), and a subset beyond the scratch limit is refused with
an error that names the limit.
"""
import numpy as np
import pytest

from fourphonon_amd import crystal
from fourphonon_amd import runtime as rt

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def _equal(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y)), (x, y)


def test_noalloc_counts_match_the_allocating_entry_point():
    case = crystal.config_case(1, ngrid=(8, 8, 8))
    banner("no-alloc NP_driver_4ph_modes, diamond 8x8x8 shell 1")
    ctx = rt.Context(case)
    try:
        nm = ctx.nmodes
        first = np.arange(1, nm + 1, 2, dtype=np.int32)          # every other mode
        second = np.array([2, 9, 10, 33, nm - 1], np.int32)      # a smaller subset next
        for modes in (first, second):
            banner(f"no-alloc call with {len(modes)} modes")
            got = ctx.NP_driver_4ph_modes_noalloc(modes)
            want = ctx.NP_driver_4ph_modes(modes)
            _equal(got, want)
            assert sum(int(n.sum()) for n in got[0::2]) > 0
        # and the full-mesh driver's values for the same modes
        full = ctx.NP_driver_4ph()
        got = ctx.NP_driver_4ph_modes_noalloc(second)
        for ch in range(3):
            assert np.array_equal(got[2 * ch], full[2 * ch][second - 1])
            assert np.array_equal(got[2 * ch + 1], full[2 * ch + 1].ravel(order="F")[second - 1])
    finally:
        ctx.close()


def test_noalloc_rejects_a_subset_beyond_the_scratch_limit():
    case = crystal.config_case(2)                                 # 16^3: 870 modes need ~128 MB
    banner("no-alloc size limit, diamond 16x16x16")
    ctx = rt.Context(case)
    try:
        modes = np.arange(1, ctx.nmodes + 1, dtype=np.int32)
        with pytest.raises(rt.FP4Error) as e:
            ctx.NP_driver_4ph_modes_noalloc(modes)
        assert e.value.code == 1 and "FP4_NOALLOC_SCRATCH_BYTES" in str(e.value)
        # a subset inside the limit still runs after the refusal
        small = np.array([1, 100, 500], np.int32)
        _equal(ctx.NP_driver_4ph_modes_noalloc(small), ctx.NP_driver_4ph_modes(small))
    finally:
        ctx.close()
