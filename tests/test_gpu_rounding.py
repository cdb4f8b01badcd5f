"""The light loop's square root and reciprocal (inv_dist in shb_common.h:
the hardware estimates corrected as the IEEE sequences correct them, without
their range scaling) give the same bits as sqrtf and 1.0f / x for every float
of the ranges they are used on: d2 in [2^-96, 2^126] and distances in
[2^-48, 2^63]. Every bake's light distances (light.h:77-93, 122-143,
161-188) go through them, so the oracle's IEEE arithmetic stays the
reference for each bit."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_corrected_sqrt_and_reciprocal_equal_ieee_on_every_float_in_range():
    lib = ctypes.CDLL(os.path.join(ROOT, "sh-baker_amd", "shb", "libshb_probe.so"))
    out = (ctypes.c_ulonglong * 6)()
    print("exhaustive sqrt / reciprocal check", flush=True)
    assert lib.shb_probe_rounding(out) == 0
    sqrt_bad, sqrt_n, rcp_bad, rcp_n, sqrt_first, rcp_first = list(out)
    print(f"  sqrt: {sqrt_n} floats, {sqrt_bad} differ; reciprocal: {rcp_n} floats, "
          f"{rcp_bad} differ")
    # every positive float from 2^-96 to 2^126 and from 2^-48 to 2^63
    assert sqrt_n == (126 + 96) * 2 ** 23 + 1
    assert rcp_n == (63 + 48) * 2 ** 23 + 1
    assert sqrt_bad == 0, hex(sqrt_first)
    assert rcp_bad == 0, hex(rcp_first)
