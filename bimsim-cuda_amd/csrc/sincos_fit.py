"""Coefficients of sincos_fast.h's polynomials (run once; needs mpmath).

sin r = r + r^3 P(z) and cos r = 1 - z/2 + z^2 Q(z), z = r^2, |r| <= pi/4:
P of degree 5 and Q of degree 4 fitted by least squares to the relative
error of sin and cos on 400 Chebyshev nodes of z in [0, (pi/4)^2], in
200-bit arithmetic, then rounded to doubles.  Prints the coefficients and
the largest relative error of the rounded polynomials (exact arithmetic)
on a fine grid of r.
"""
import mpmath as mp

mp.mp.prec = 200
ZMAX = (mp.pi / 4) ** 2


def fit(f, deg, n=400):
    nodes = [ZMAX * (1 - mp.cos(mp.pi * (k + 0.5) / n)) / 2 for k in range(n)]
    A = mp.matrix(n, deg + 1)
    b = mp.matrix(n, 1)
    for i, z in enumerate(nodes):
        val, w = f(z)
        for j in range(deg + 1):
            A[i, j] = z ** j * w
        b[i] = val * w
    x = mp.lu_solve(A.T * A, A.T * b)
    return [float(x[j]) for j in range(deg + 1)]


def sin_part(z):
    """(sin r - r) / r^3 and the weight r^3 / sin r of a relative error."""
    r = mp.sqrt(z)
    return (mp.sin(r) - r) / r ** 3, r ** 3 / mp.sin(r)


def cos_part(z):
    """(cos r - 1 + z/2) / z^2 and the weight z^2 / cos r."""
    r = mp.sqrt(z)
    return (mp.cos(r) - 1 + z / 2) / z ** 2, z ** 2 / mp.cos(r)


def horner(coef, z):
    acc = mp.mpf(coef[-1])
    for c in reversed(coef[:-1]):
        acc = acc * z + mp.mpf(c)
    return acc


if __name__ == "__main__":
    S = fit(sin_part, 5)
    C = fit(cos_part, 4)
    print("kS =", ", ".join("%.17g" % c for c in S))
    print("kC =", ", ".join("%.17g" % c for c in C))
    es = ec = mp.mpf(0)
    for k in range(1, 2001):
        r = (mp.pi / 4) * k / 2000
        z = r * r
        es = max(es, abs((r + r ** 3 * horner(S, z)) / mp.sin(r) - 1))
        ec = max(ec, abs((1 - z / 2 + z * z * horner(C, z)) / mp.cos(r) - 1))
    print("largest relative error: sin %s, cos %s" % (mp.nstr(es, 3), mp.nstr(ec, 3)))
