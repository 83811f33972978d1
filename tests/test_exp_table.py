"""The exponential table of the batched integrators (cubnm_amd/csrc/exp_table.hpp).

Checks the header's constants against exact arithmetic, on the CPU: every table
entry is 2^(j/256) correctly rounded to double (60-digit decimal arithmetic),
the multiplier is 256/ln 2, the Cody-Waite head carries enough trailing zero
bits for k * head to be exact over the arguments that do not overflow or
underflow, head + tail is ln 2 / 256 to well below an ulp, and the degree-4
Taylor polynomial's truncation error over |r| <= ln2/512 is below half an ulp
of 1.  The kernel's use of the table is covered by the GPU parity tests.
"""
import math
import re
from decimal import Decimal, getcontext
from pathlib import Path

import pytest

HEADER = Path(__file__).resolve().parents[1] / "cubnm_amd" / "csrc" / "exp_table.hpp"


@pytest.fixture(scope="module")
def src():
    return HEADER.read_text()


def _table(src):
    body = src[src.index("k_exp2_256ths[kExpTableSize] = {"):]
    body = body[: body.index("};")]
    return [float.fromhex(h) for h in re.findall(r"0x[0-9a-f.]+p[+-]\d+", body)]


def test_table_entries_are_correctly_rounded(src):
    tab = _table(src)
    assert len(tab) == 256
    getcontext().prec = 60
    ln2 = Decimal(2).ln()
    for j, v in enumerate(tab):
        exact = (Decimal(j) * ln2 / 256).exp()
        assert v == float(exact), j
        # correct rounding: no other double is nearer
        err = abs(Decimal(v) - exact)
        assert err <= abs(Decimal(math.nextafter(v, math.inf)) - exact)
        assert err <= abs(Decimal(math.nextafter(v, -math.inf)) - exact)


def test_reduction_constants(src):
    mult = float(re.search(r"x \* ([0-9.]+)\)", src).group(1))
    assert mult == pytest.approx(256 / math.log(2), rel=1e-15)
    head = float(re.search(r"fma\(-k, ([0-9.e+-]+) / 256, x\)", src).group(1))
    tail = float(re.search(r"fma\(-k, ([0-9.e+-]+) / 256, r\)", src).group(1))
    # head has trailing zero bits: k * head / 256 is exact for |k| < 2^21
    # (|x| < 5,678, beyond which exp overflows or underflows anyway)
    mant = int(head.hex().split(".")[1].split("p")[0], 16)
    trailing = (mant & -mant).bit_length() - 1
    assert trailing >= 20
    getcontext().prec = 60
    ln2 = Decimal(2).ln()
    assert abs(Decimal(head) + Decimal(tail) - ln2) < Decimal(2) ** -80


def test_polynomial_degree_suffices():
    r = math.log(2) / 512
    # exp(r) - 1 = r + r^2/2 + r^3/6 + r^4/24 + (r^5/120 + ...): truncation after r^4
    trunc = sum(r ** n / math.factorial(n) for n in range(5, 12))
    assert trunc < 2.0 ** -54
    # a degree-3 polynomial would not do
    assert r ** 4 / 24 > 2.0 ** -53
