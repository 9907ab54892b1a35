"""The allocation-free MRD block entry point (coresg_hip_mrd_block_noalloc)
against coresg_hip_mrd_block, through the C ABI: bit-identical rows on two
calls in a row, the second smaller, and a clear error above the fixed size."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_pkg_module

chip = load_pkg_module("coresg_hip")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu


def _case(n, d, seed):
    x = datagen.gaussian_mixture(n, d, 6, 0.3, 2.0, 0.05, seed=seed)
    stride = (d + 7) & ~7
    xp = np.zeros((n, stride), np.float32)
    xp[:, :d] = x
    rng = np.random.default_rng(seed)
    core = rng.uniform(0.0, 1.5, size=n).astype(np.float32)
    core[::5] = 0.0
    return xp, d, core


def test_noalloc_mrd_block_equals_mrd_block_twice_smaller_second():
    print("[gpu] mrd_block_noalloc: 1500x24 rows 0..1500 then 700x16 rows 100..400", flush=True)
    cap = chip.lib().coresg_hip_mrd_block_noalloc_capacity()
    assert cap == 8 << 20
    for n, d, seed, r0, r1 in ((1500, 24, 1, 0, 1500), (700, 16, 2, 100, 400)):
        xp, dim, core = _case(n, d, seed)
        want = chip.mrd_block(xp, dim, core, r0, r1)
        got = chip.mrd_block_noalloc(xp, dim, core, r0, r1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        # row-at-a-time, as Prim asks for rows
        for r in (r0, (r0 + r1) // 2, r1 - 1):
            one = chip.mrd_block_noalloc(xp, dim, core, r, r + 1)
            assert np.array_equal(one.view(np.uint32), want[r - r0:r - r0 + 1].view(np.uint32))


def test_noalloc_mrd_block_rejects_sizes_above_the_scratch():
    print("[gpu] mrd_block_noalloc: oversize request rejected", flush=True)
    n = 3000  # 3000 x 3000 rows alone need 9M floats > 8M
    xp, dim, core = _case(n, 8, 3)
    with pytest.raises(chip.CoresgHipError, match="fixed scratch"):
        chip.mrd_block_noalloc(xp, dim, core, 0, n)
    # still usable afterwards
    small = chip.mrd_block_noalloc(xp[:200], dim, core[:200], 0, 200)
    want = chip.mrd_block(xp[:200], dim, core[:200], 0, 200)
    assert np.array_equal(small.view(np.uint32), want.view(np.uint32))
