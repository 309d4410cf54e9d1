"""Coefficients of the FP64 arctangent kernel used by K1's phase (atan_f64.h).

Method (range reduction by the tan(pi/8) identity, polynomial correction):
with 0 <= a <= b, atan(a/b) = atan(t) for t = a/b <= tan(pi/8), else
pi/4 + atan(t) for t = (a - b)/(a + b), so |t| <= tan(pi/8) and
u = t^2 <= U = tan^2(pi/8) = 3 - 2 sqrt(2).  On that range
    atan(t) = t + t u P(u),   P(u) ~ f(u) = (atan(sqrt u)/sqrt u - 1)/u
                                         = sum_k (-1)^(k+1) u^k / (2k + 3).
P is the Chebyshev interpolant of f at N nodes on [0, U] (f evaluated from
its series with 50 significant digits), expanded to monomials in u and
rounded to double.  Prints the coefficients c0..c(N-1) (P = sum c_k u^k) and
the interpolation error bound |f - P| on a dense grid.
"""
from decimal import Decimal, getcontext

getcontext().prec = 60
N = 11


def f(u):
    s, k, term = Decimal(0), 0, Decimal(1)
    while True:
        t = term / (2 * k + 3)
        s += -t if k % 2 == 0 else t
        if t < Decimal(10) ** -55:
            return s
        term *= u
        k += 1


def cheb_poly(n):
    """Monomial coefficients of T_0..T_{n-1} (in x)."""
    T = [[Decimal(1)], [Decimal(0), Decimal(1)]]
    while len(T) < n:
        a, b = T[-1], T[-2]
        c = [Decimal(0)] + [2 * v for v in a]
        for i, v in enumerate(b):
            c[i] -= v
        T.append(c)
    return T[:n]


def main():
    two = Decimal(2)
    U = 3 - two * two.sqrt()
    pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494")
    nodes = [(Decimal(j) + Decimal("0.5")) * pi / N for j in range(N)]

    def cos(x):
        s, term, k = Decimal(0), Decimal(1), 0
        while abs(term) > Decimal(10) ** -58:
            s += term
            k += 2
            term *= -x * x / (k * (k - 1))
        return s

    xs = [cos(x) for x in nodes]
    fs = [f(U * (x + 1) / 2) for x in xs]
    T = cheb_poly(N)
    a = []
    for k in range(N):
        tk = [sum(c * x ** i for i, c in enumerate(T[k])) for x in xs]
        a.append((Decimal(2) if k else Decimal(1)) / N * sum(fv * t for fv, t in zip(fs, tk)))
    # sum_k a_k T_k(x) with x = 2u/U - 1 -> monomials in u
    px = [Decimal(0)] * N
    for k in range(N):
        for i, c in enumerate(T[k]):
            px[i] += a[k] * c
    pu = [Decimal(0)] * N
    # x^i = (2u/U - 1)^i
    from math import comb
    for i, c in enumerate(px):
        for j in range(i + 1):
            pu[j] += c * comb(i, j) * (2 / U) ** j * (-1) ** (i - j)
    coef = [float(c) for c in pu]
    err = Decimal(0)
    for g in range(2001):
        u = U * g / 2000
        p = Decimal(0)
        for c in reversed(coef):
            p = p * u + Decimal(c)
        err = max(err, abs(p - f(u)))
    print("N", N, "max |f - P| on [0, U] with double coefficients:", f"{float(err):.3e}")
    print("t u |f - P| bound relative to atan(t):", f"{float(err * U):.3e}")
    for k, c in enumerate(coef):
        print(f"    {c!r},  // c{k}")


if __name__ == "__main__":
    main()
