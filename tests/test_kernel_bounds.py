"""Numeric bounds the HIP kernels rely on, checked on the CPU.

The self-guided A/B step runs p * s and (256 - a2) * oneOverN * b on the
24-bit multiplier and in 32 bits (lr.hpp, lr_ab), where the reference
(loop_restoration.rs:640-661) computes in 64 bits with a saturating
multiply, and packs A and B into one word whose neighbours are summed
before the fields are split (lr_pack_ab); and the packed CDEF filter (cdef.hpp, cdef_filter_plane_pk) skips
the final clip to the taps' min/max when a unit has only primary or only
secondary taps (cdef.rs:592-640 always clips).  Both are exact only under
the bounds below."""
import json
import os
import random

import pytest

from conftest import ROOT

FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "spec_tables.json")))


def _sgr_s(r, e):
    """s = (2^20 + n^2 e / 2) / (n^2 e) (loop_restoration.rs:612-616)."""
    n = (2 * r + 1) ** 2
    n2e = n * n * e
    return ((1 << 20) + n2e // 2) // max(n2e, 1)


def _round2(x, sh):
    return (x + ((1 << sh) >> 1)) >> sh


def _worst_p(n, bd):
    """Largest p = Round2(a) * n - Round2(b)^2 over windows of n samples
    taking a few values near 0 and near the maximum (the variance is
    largest for two-valued windows at the extremes)."""
    mx = (1 << bd) - 1
    lows = (0, 1, 2, 3)
    highs = (mx, mx - 1, mx - 2, mx - 3)
    best = 0
    for k in range(n + 1):
        for lo in lows:
            for hi in highs:
                a = k * hi * hi + (n - k) * lo * lo
                b = k * hi + (n - k) * lo
                p = _round2(a, 2 * (bd - 8)) * n - _round2(b, bd - 8) ** 2
                best = max(best, p)
    return best


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_self_guided_products_fit_32_bits(bd):
    params = FIX["SGR_PARAMS"]
    assert len(params) == 64
    for st in range(16):
        for ps in range(2):
            r, e = params[4 * st + 2 * ps], params[4 * st + 2 * ps + 1]
            if r == 0:
                continue
            n = (2 * r + 1) ** 2
            s = _sgr_s(r, e)
            p = _worst_p(n, bd)
            # p on the 24-bit multiplier, p * s below 2^32 (no wrap, so the
            # reference's saturating 64-bit product is the same number)
            assert p < 1 << 24, (st, ps, bd, p)
            assert p * s < 1 << 32, (st, ps, bd, p, s)
            assert s < 1 << 12
            # the operands of the other products
            ar = _round2(n * ((1 << bd) - 1) ** 2, 2 * (bd - 8))
            dr = _round2(n * ((1 << bd) - 1), bd - 8)
            assert ar < 1 << 21 and dr <= 25 * 256
            one_over_n = ((1 << 12) + n // 2) // n
            b = n * ((1 << bd) - 1)
            assert b < 1 << 17
            assert 255 * one_over_n * b < 1 << 32
            # the packed A/B word (lr_pack_ab): A <= 256 at bit 22, B below;
            # sums of three words must not carry from B into A
            big_b = _round2(255 * one_over_n * b, 12)
            assert big_b < 1 << 20
            assert 3 * big_b < 1 << 22 and 3 * 256 < 1 << 10 and 256 << 22 < 1 << 32


def _constrain(diff, threshold, damping):
    """constrain (cdef.rs:592-606)."""
    if not threshold:
        return 0
    shift = max(0, damping - (threshold.bit_length() - 1))
    mag = min(abs(diff), max(0, threshold - (abs(diff) >> shift)))
    return mag if diff >= 0 else -mag


@pytest.mark.parametrize("kind", ["primary", "secondary"])
def test_cdef_one_kind_of_taps_stays_inside_the_taps(kind):
    """With only primary taps (weights 4, 2 or 3, 3 on both sides of the
    direction) or only secondary taps (2, 1 on four taps each), the weights
    sum to 12 < 16, so x + ((8 + sum - (sum < 0)) >> 4) lies between the
    smallest and the largest of x and its taps: the clip is the identity."""
    rng = random.Random(0xCDEF0 + (kind == "secondary"))
    for _ in range(20000):
        bd = rng.choice([8, 10, 12])
        mx = (1 << bd) - 1
        x = rng.randint(0, mx)
        if kind == "primary":
            w = rng.choice([(4, 2), (3, 3)])
            weights = [w[0], w[0], w[1], w[1]]
            strength = rng.randint(1, 15) << (bd - 8)
        else:
            weights = [2, 2, 2, 2, 1, 1, 1, 1]
            strength = rng.choice([1, 2, 4]) << (bd - 8)
        damping = rng.randint(3, 6) + (bd - 8)
        # taps: far, near and equal values, so clamped and unclamped
        # differences both occur
        taps = [rng.choice([rng.randint(0, mx), min(mx, max(0, x + rng.randint(-40, 40))), x])
                for _ in weights]
        total = sum(wt * _constrain(v - x, strength, damping) for wt, v in zip(weights, taps))
        y = x + ((8 + total - (total < 0)) >> 4)
        assert min(taps + [x]) <= y <= max(taps + [x]), (kind, x, taps, y)
