"""Table-driven exp of the FP64 neighbour pass (exp_reduce / exp_finish in
iwf_mfree_gpu_3d_amd/csrc/k_pair.hip).

The PSE heat sum (src/interactions_gpu.cu:124-147, kernel src/kernels.cuh:71-92) evaluates
exp(-r^2 / h^2) for every candidate pair; the GPU pass uses Tang's table-driven method instead of
the general library routine.  Pinned here on the CPU:
- its constants (read from the kernel source) are the correctly rounded values of 2^(j/64) and
  of the remainder 2^(j/64) - lead, of 64/ln2 and of the two-part ln2/64, against a 60-digit
  decimal evaluation (and 2^(1/2) against the correctly rounded square root);
- the method, carried out with the same IEEE double operations (each fused multiply-add exact
  through rationals, rounded once), stays within 0.85 ulp of exp over the arguments the pass
  produces, the underflow range and its edges.
The device code itself is pinned by the FP64 one-step tests (relative L2 1e-12 against the
oracle's libm exp).
"""
import math
import random
import re
from decimal import Decimal, getcontext
from fractions import Fraction
from pathlib import Path

SRC = Path(__file__).resolve().parents[1] / "iwf_mfree_gpu_3d_amd" / "csrc" / "k_pair.hip"
getcontext().prec = 60
LN2 = Decimal(2).ln()


def _constants():
    block = SRC.read_text().split("// BEGIN EXP TABLE")[1].split("// END EXP TABLE")[0]
    named = {k: float.fromhex(v)
             for k, v in re.findall(r"constexpr double (\w+) = (0x[0-9a-f.]+p[+-]\d+);", block)}
    v = [float.fromhex(x) for x in re.findall(r"-?0x[0-9a-f.]+p[+-]\d+", block.split("= {")[1])]
    return named, list(zip(v[0::2], v[1::2]))


def _rn(d: Decimal) -> float:
    return float(str(d))  # CPython's string conversion rounds correctly


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _exp_nonpos(a, C, tab):
    a = max(a, -746.0)
    n = float(round(a * C["EXP_INV_LN2_64"]))  # round(): ties to even, as rint
    r = _fma(-n, C["EXP_LN2_64_HI"], a)
    r = _fma(-n, C["EXP_LN2_64_LO"], r)
    ni = int(n)
    p = _fma(r, 1.0 / 120.0, 1.0 / 24.0)
    p = _fma(p, r, 1.0 / 6.0)
    p = _fma(p, r, 0.5)
    p = _fma(p, r, 1.0)
    p = p * r
    lead, trail = tab[ni & 63]
    return math.ldexp(lead + _fma(lead, p, trail), ni >> 6)


def test_constants_correctly_rounded():
    C, tab = _constants()
    assert len(tab) == 64
    for j, (lead, trail) in enumerate(tab):
        e = (LN2 * j / 64).exp()
        assert lead == _rn(e), j
        assert trail == _rn(e - Decimal(lead)), j
    assert tab[0] == (1.0, 0.0) and tab[32][0] == math.sqrt(2.0)
    assert C["EXP_INV_LN2_64"] == _rn(64 / LN2)
    assert C["EXP_LN2_64_HI"] == _rn(LN2 / 64)
    assert C["EXP_LN2_64_LO"] == _rn(LN2 / 64 - Decimal(C["EXP_LN2_64_HI"]))


def test_method_within_0p85_ulp():
    C, tab = _constants()
    rng = random.Random(5)
    args = [0.0, -0.0, -1e-300, -1e-17, -math.log(2) / 128, -math.log(2) / 64, -1.0, -4.0,
            -27.0, -708.0, -745.1, -745.2, -746.0, -800.0, -1e12]
    args += [-rng.uniform(0.0, 30.0) for _ in range(2500)]       # r^2 / h^2 inside a stencil
    args += [-rng.uniform(30.0, 746.0) for _ in range(300)]      # far candidates, underflow
    # both sides of the reduction boundaries (n changes, |r| = ln2/128)
    for k in range(1, 400, 7):
        b = -(k + 0.5) * math.log(2) / 64
        args += [math.nextafter(b, 0.0), b, math.nextafter(b, -1.0)]
    assert _exp_nonpos(0.0, C, tab) == 1.0 and _exp_nonpos(-0.0, C, tab) == 1.0
    worst = 0.0
    for a in args:
        y = _exp_nonpos(a, C, tab)
        exact = Decimal(a).exp() if a > -800 else Decimal(0)
        ref = _rn(exact)
        if ref == 0.0:
            assert y <= 5e-324, (a, y)
            continue
        err = abs(Decimal(y) - exact) / Decimal(math.ulp(ref))
        worst = max(worst, float(err))
        # bound: final rounding 0.5 ulp + truncation r^6/720 (up to 0.31 ulp for results near 2)
        assert err <= Decimal("0.85"), (a, y, ref, float(err))
        assert abs(y - math.exp(a)) <= 2 * math.ulp(ref), (a, y, math.exp(a))
    assert worst > 0.3, worst  # the samples reach the rounding regime

