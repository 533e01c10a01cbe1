"""The fused FP64 pass computes K = exp(-gamma d2) with its own branch-free
exp (exp_kernel_f64 in csrc/rbf_f64.hip: Cody-Waite reduction by ln 2 split
into a 32-bit head and a tail, Taylor polynomial of degree 13, ldexp).

This CPU test reads the constants from the source, evaluates the same steps
with exactly rounded fused multiply-adds (fractions), and checks the result
against math.exp: within 1 ulp over [-746, 0], exact at 0, zero below the
subnormal range.  It also checks each constant against its definition (the
head has 32 significant bits, head + tail = ln 2 to double precision, the
coefficients are 1/k! rounded to double), so a mistyped digit fails here."""
import math
import random
import re
from fractions import Fraction
from pathlib import Path

SRC = Path(__file__).resolve().parents[1] / "darray_optimizer_amd" / "csrc" / "rbf_f64.hip"


def constants():
    body = SRC.read_text()
    body = body[body.index("double exp_kernel_f64(double x)"):]
    body = body[:body.index("return ldexp(p, (int)k);")]
    hexes = re.findall(r"(-?0x[0-9a-fA-F.]+p-?\d+)", body)
    hi, lo = float.fromhex(hexes[0]), float.fromhex(hexes[1])
    log2e = float(re.search(r"rint\(x \* ([0-9.e+-]+)\)", body).group(1))
    coef = [float(v) for v in re.findall(r"p = (?:fma\(p, r, )?([0-9.]+e-?\d+|0\.5|1\.0)", body)]
    return hi, lo, log2e, coef  # coef from degree 13 down to 0


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def device_exp(x, hi, lo, log2e, coef):
    x = max(x, -746.0)
    k = float(round(x * log2e))  # rint: ties to even, like v_rndne_f64
    r = fma(-k, hi, x)
    r = fma(-k, lo, r)
    p = coef[0]
    for c in coef[1:]:
        p = fma(p, r, c)
    return math.ldexp(p, int(k))


def test_constants_match_their_definitions():
    hi, lo, log2e, coef = constants()
    assert len(coef) == 14
    for j, c in enumerate(coef):
        assert c == float(Fraction(1, math.factorial(13 - j)))
    m, e = math.frexp(hi)
    assert (Fraction(m) * 2**32).denominator == 1  # 32 significant bits
    assert abs(hi + lo - math.log(2)) <= math.ulp(math.log(2))
    assert log2e == 1.0 / math.log(2) or abs(log2e - 1.0 / math.log(2)) <= math.ulp(log2e)


def test_within_one_ulp_of_libm():
    hi, lo, log2e, coef = constants()
    rng = random.Random(3)
    xs = [0.0, -0.0, -1e-300, -1e-12, -0.5, -math.log(2) / 2, -1.0, -20.0, -700.0, -708.39,
          -709.0, -744.0, -745.1, -746.0, -1000.0]
    xs += [-rng.random() * rng.choice([1e-6, 1e-2, 1.0, 10.0, 100.0, 745.0]) for _ in range(3000)]
    for x in xs:
        a, b = device_exp(x, hi, lo, log2e, coef), math.exp(x)
        if b == 0.0:
            assert a <= 5e-324, (x, a)
            continue
        assert abs(a - b) <= math.ulp(b), (x, a, b)
    assert device_exp(0.0, hi, lo, log2e, coef) == 1.0
