"""Device integer helpers against exact arithmetic: the batched claim step's
multiply-shift divisions and the texel index conversions of filter_level
(pathtrace.wgsl:335-380), at their extremes."""
import numpy as np
import pytest

import oracle
import oxpt
from test_oracle_pins import UDIV_DIVISORS, _texel_ref, _udiv_numerators, texel_cases

pytestmark = pytest.mark.gpu


def test_device_udiv_is_exact(gpu):
    num, den = [], []
    for d in UDIV_DIVISORS:
        for n in _udiv_numerators(d):
            num.append(n)
            den.append(d)
    rng = np.random.default_rng(3)
    # Random pairs, with divisors of the claim step's range (64 x spp, tiles_x).
    rn = rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64)
    rd = np.concatenate([rng.integers(1, 1 << 14, 1 << 19, dtype=np.uint64),
                         rng.integers(1, 1 << 32, 1 << 19, dtype=np.uint64)])
    num = np.concatenate([np.array(num, np.uint64), rn]).astype(np.uint32)
    den = np.concatenate([np.array(den, np.uint64), rd]).astype(np.uint32)
    got = gpu.debug_udiv(num, den)
    want = (num.astype(np.uint64) // den.astype(np.uint64)).astype(np.uint32)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(int(num[i]), int(den[i]), int(got[i]), int(want[i])) for i in bad[:5]]


def test_device_texel_index_matches_the_reference_rules(gpu):
    """|u| up to the 1e6 clamp on levels up to 65535 texels (products beyond
    the i32 range saturate), non-finite u, all three wrap modes: the device
    equals the reference's rules and the oracle."""
    cases = texel_cases()
    rng = np.random.default_rng(9)
    for _ in range(20000):
        cases.append((float(rng.uniform(-1e6, 1e6)), int(rng.integers(1, 65536)), int(rng.integers(0, 3))))
    u = np.array([c[0] for c in cases], np.float32)
    size = np.array([c[1] for c in cases], np.int32)
    mode = np.array([c[2] for c in cases], np.uint32)
    got = gpu.debug_texel_index(u, size, mode)
    for k, (uu, n, m) in enumerate(cases):
        want = _texel_ref(uu, n, m)
        assert tuple(int(v) for v in got[k]) == want, (uu, n, m, got[k], want)
        assert all(0 <= int(v) < n for v in got[k])
        if k < 3000:
            assert oracle.texel_index(uu, n, m) == want
