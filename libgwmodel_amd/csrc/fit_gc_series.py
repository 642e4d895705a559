"""Coefficients of gc_d2_near_f64 (gwm_device.h): minimax polynomials on
S in [0, 1/8] of asin(sqrt(S))^2 / S (degree 10) and of sinc(2 asin(sqrt(S)))
(degree 7), from mpmath's Chebyshev fit, printed highest degree first.
Run: python3 fit_gc_series.py"""
import mpmath as mp

mp.mp.dps = 40
SMAX = mp.mpf(1) / 8


def q(S):
    return mp.mpf(1) if S == 0 else mp.asin(mp.sqrt(S)) ** 2 / S


def sinc2w(S):
    x = 2 * mp.sqrt(S * q(S))
    return mp.mpf(1) if x == 0 else mp.sin(x) / x


for name, fn, deg in (("w^2/S", q, 10), ("R", sinc2w, 7)):
    poly, err = mp.chebyfit(fn, [0, SMAX], deg + 1, error=True)
    print("%s: max error %s" % (name, mp.nstr(err, 3)))
    print("  " + ", ".join(repr(float(c)) for c in poly))
