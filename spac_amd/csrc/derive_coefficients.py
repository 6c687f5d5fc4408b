"""Derive the polynomial coefficients of spac_math.h's exp, log, sinpi and sincospi.

Each function is reduced to a small interval and the remainder after its exact leading terms is
approximated by a Chebyshev interpolant (mpmath.chebyfit at 60 digits, close to the minimax
polynomial), the shortest whose approximation error is below `ULP_BUDGET` units in the last place
of the result; the coefficients are then rounded to the nearest double.  The evaluation rounding
(a Horner chain in FMAs) dominates the total error, which the math tests bound against numpy.

    python3 spac_amd/csrc/derive_coefficients.py      # prints the C literals, highest degree first
"""
import mpmath as mp

mp.mp.dps = 60
ULP = mp.mpf(2) ** -53
ULP_BUDGET = mp.mpf("0.02")


def _fit(name, g, lo, hi, rel_err, max_terms=16):
    """Shortest Chebyshev fit of g on [lo, hi] whose relative error rel_err(poly, w) stays below
    the budget on a dense grid (exact coefficients); also reports the error once the
    coefficients are rounded to doubles, as they are used (the rounding of the leading
    coefficient, e.g. -pi^3/6, is part of every such polynomial, Taylor or fitted)."""
    grid = [lo + (hi - lo) * i / 2000 for i in range(1, 2001)]
    for n in range(3, max_terms + 1):
        poly, _ = mp.chebyfit(g, [lo, hi], n, error=True)
        err = max(abs(rel_err(poly, w)) for w in grid)
        if err < ULP_BUDGET * ULP:
            rounded = [mp.mpf(float(c)) for c in poly]
            err_r = max(abs(rel_err(rounded, w)) for w in grid)
            print(f"{name}: {n} coefficients, approximation error {mp.nstr(err / ULP, 3)} ulp "
                  f"({mp.nstr(err_r / ULP, 3)} ulp with double coefficients)")
            for c in rounded:
                print(f"    {float(c)!r}")
            return rounded
    raise RuntimeError(name)


def exp_poly():
    # e^r = 1 + r + r^2 Q(r) on |r| <= ln2 / 2 (with margin for the rounding of k)
    L = mp.log(2) / 2 * (1 + mp.mpf("1e-6"))
    Q = lambda r: mp.mpf(1) / 2 + r / 6 if abs(r) < mp.mpf("1e-30") else (mp.exp(r) - 1 - r) / r ** 2
    return _fit("exp Q(r)", Q, -L, L, lambda p, r: r * r * (mp.polyval(p, r) - Q(r)) / mp.exp(r))


def log_poly():
    # log m = 2 f + f^3 R(f^2), f = (m - 1) / (m + 1), m in [sqrt(2)/2, sqrt(2)]
    smax = ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2 * (1 + mp.mpf("1e-9"))
    R = lambda s: mp.mpf(2) / 3 if s < mp.mpf("1e-40") else (2 * mp.atanh(mp.sqrt(s)) - 2 * mp.sqrt(s)) / s ** 1.5
    return _fit("log R(s)", R, mp.mpf("1e-40"), smax,
                lambda p, s: s ** 1.5 * (mp.polyval(p, s) - R(s)) / (2 * mp.atanh(mp.sqrt(s))))


def sinpi_half_poly():
    # sin(pi t) = pi t + t^3 P(t^2) on |t| <= 1/2
    P = lambda w: -mp.pi ** 3 / 6 if w < mp.mpf("1e-40") else (mp.sin(mp.pi * mp.sqrt(w)) - mp.pi * mp.sqrt(w)) / w ** 1.5
    return _fit("sinpi_half P(t^2)", P, mp.mpf("1e-40"), mp.mpf(1) / 4,
                lambda p, w: w ** 1.5 * (mp.polyval(p, w) - P(w)) / mp.sin(mp.pi * mp.sqrt(w)))


def sincospi_polys():
    # sin(pi r) = pi r + r^3 S(r^2), cos(pi r) = 1 + r^2 C(r^2) on |r| <= 1/4
    S = lambda w: -mp.pi ** 3 / 6 if w < mp.mpf("1e-40") else (mp.sin(mp.pi * mp.sqrt(w)) - mp.pi * mp.sqrt(w)) / w ** 1.5
    C = lambda w: -mp.pi ** 2 / 2 if w < mp.mpf("1e-40") else (mp.cos(mp.pi * mp.sqrt(w)) - 1) / w
    s = _fit("sincospi S(r^2)", S, mp.mpf("1e-40"), mp.mpf(1) / 16,
             lambda p, w: w ** 1.5 * (mp.polyval(p, w) - S(w)) / mp.sin(mp.pi * mp.sqrt(w)))
    c = _fit("sincospi C(r^2)", C, mp.mpf("1e-40"), mp.mpf(1) / 16,
             lambda p, w: w * (mp.polyval(p, w) - C(w)) / mp.cos(mp.pi * mp.sqrt(w)))
    return s, c


if __name__ == "__main__":
    exp_poly()
    log_poly()
    sinpi_half_poly()
    sincospi_polys()
