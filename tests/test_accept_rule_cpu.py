"""The accept test the kernels apply without opacities: a pair is kept iff
s = sigma * log2(e) lies in [+0, kAcceptS] (bit patterns compared as unsigned
integers), with kAcceptS the largest FP32 value not above log2(255).  The
reference keeps a pair unless sigma < 0 or min(exp(-sigma), clamp) < 1/255
(gsplat2d/gsplat2d/cuda/csrc/forward.cu:254-263, backward.cu:105-111)."""
import math
import os
import re
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ACCEPT_BITS = 0x40FFD1BE


def _bits(x):
    return int(np.array([x], np.float32).view(np.uint32)[0])


def _f(bits):
    return float(np.array([bits], np.uint32).view(np.float32)[0])


def test_constant_is_largest_float_not_above_log2_255():
    s = _f(ACCEPT_BITS)
    nxt = _f(ACCEPT_BITS + 1)
    # 2^s <= 255 < 2^next: s is the largest FP32 value not above log2(255)
    assert s < math.log2(255.0) < nxt
    assert 2.0 ** s <= 255.0 < 2.0 ** nxt
    assert Fraction(2) ** 0 <= Fraction(255) / Fraction(2.0 ** s)
    # the cut in alpha: exp2(-s) * 255 just above 1, the next value just below
    above = 255.0 * 2.0 ** (-s) - 1.0
    below = 1.0 - 255.0 * 2.0 ** (-nxt)
    assert 0.0 < above < 1e-7 and 0.0 < below < 2.5e-7
    # both far inside the oracle's ambiguity band |255 alpha - 1| < 1e-5
    assert above < 1e-5 / 50 and below < 1e-5 / 40


def _accept_rule(s):
    return np.array([s], np.float32).view(np.uint32)[0] <= ACCEPT_BITS


def _reference_rule(s):
    # FP64 exponential, as the oracle: keep unless sigma < 0 or alpha < 1/255
    if not (s == s):
        return True  # fminf(NaN, clamp) = clamp: kept at full alpha
    sigma = float(s) / math.log2(math.e)
    if sigma < 0:
        return False
    alpha = min(math.exp(-sigma), 0.999)
    return not (alpha < 1.0 / 255.0)


def test_rule_agrees_with_reference_outside_the_ambiguity_band():
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.uniform(-1, 12, 200000),
        _f(ACCEPT_BITS) + np.arange(-4000, 4000) * 2.0 ** -21,
        [0.0, 1e-30, -1e-30, 7.99, 8.0, np.inf, -np.inf]]).astype(np.float32)
    for v in vals:
        v = np.float32(v)
        ref = _reference_rule(v)
        got = _accept_rule(v)
        if got != ref:
            # only where the oracle reports the pair as ambiguous
            sigma = float(v) / math.log2(math.e)
            alpha = min(math.exp(-sigma), 0.999) if sigma >= 0 else float("inf")
            assert abs(alpha * 255.0 - 1.0) < 1e-5 or abs(sigma) < 1e-12, (float(v), ref, got)
    # the two documented differences beyond rounding
    assert _reference_rule(np.float32(np.nan)) and not _accept_rule(np.float32(np.nan))
    assert _reference_rule(np.float32(-0.0)) and not _accept_rule(np.float32(-0.0))


def test_device_header_carries_the_constant():
    text = open(os.path.join(ROOT, "lig_amd", "csrc", "lig_device.h")).read()
    m = re.search(r"kAcceptSBits\s*=\s*(0x[0-9a-fA-F]+)u", text)
    assert m and int(m.group(1), 16) == ACCEPT_BITS
