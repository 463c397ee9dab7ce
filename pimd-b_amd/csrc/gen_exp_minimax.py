"""Coefficients of exp_minimax (device_common.hpp): the degree-11 polynomial
of best relative approximation to e^r on |r| <= ln(2)/2, by the Remez
exchange algorithm in 50-digit arithmetic (mpmath).  Prints the coefficients
rounded to double and the approximation error they leave.

    python gen_exp_minimax.py [degree]
"""
import sys

import mpmath as mp


def remez(n, a):
    """Coefficients d_j of sum d_j u^j ~ e^(a u) on [-1, 1], relative error."""
    f = lambda u: mp.exp(a * u)
    us = [mp.cos(mp.pi * (n + 1 - i) / (n + 1)) for i in range(n + 2)]
    us = [u if abs(u) > 1e-30 else mp.mpf(0) for u in us]
    for _ in range(40):
        A = mp.matrix(n + 2, n + 2)
        b = mp.matrix(n + 2, 1)
        for i, u in enumerate(us):
            for j in range(n + 1):
                A[i, j] = u ** j
            A[i, n + 1] = (-1) ** i * f(u)
            b[i] = f(u)
        sol = mp.lu_solve(A, b)
        d = [sol[j] for j in range(n + 1)]
        E = sol[n + 1]
        err = lambda u: (mp.polyval(d[::-1], u) - f(u)) / f(u)
        K = 3000
        grid = [-1 + mp.mpf(2) * k / K for k in range(K + 1)]
        vals = [err(u) for u in grid]
        ext = [grid[0]]
        for k in range(1, K):
            if (vals[k] - vals[k - 1]) * (vals[k + 1] - vals[k]) < 0:
                lo, hi = grid[k - 1], grid[k + 1]
                for _ in range(60):   # ternary search for the extremum of |err|
                    m1, m2 = lo + (hi - lo) / 3, hi - (hi - lo) / 3
                    if abs(err(m1)) > abs(err(m2)):
                        hi = m2
                    else:
                        lo = m1
                ext.append((lo + hi) / 2)
        ext.append(grid[-1])
        if len(ext) != n + 2:
            raise RuntimeError("alternation lost")
        us = ext
        if abs(max(abs(err(u)) for u in ext) - abs(E)) < mp.mpf("1e-4") * abs(E):
            return d, E
    raise RuntimeError("no convergence")


def main():
    mp.mp.dps = 50
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 11
    a = mp.log(2) / 2 * (1 + mp.mpf("1e-9"))
    d, E = remez(n, a)
    c = [float(d[j] / a ** j) for j in range(n + 1)]
    for j, cj in enumerate(c):
        print("c%d = %r" % (j, cj))
    xs = [-a + 2 * a * k / 20000 for k in range(20001)]
    worst = max(abs((mp.polyval([mp.mpf(x) for x in c[::-1]], x) - mp.exp(x)) / mp.exp(x)) for x in xs)
    print("equioscillation %s; with double coefficients %s relative" % (mp.nstr(E, 5), mp.nstr(worst, 5)))


if __name__ == "__main__":
    main()
