"""Writes analytic_values.json: reference values fixed by mathematics, not by
this project's code.  The reference repository ships no golden vectors.

- central_difference_4th_order: weights of the 5-point first-derivative
  stencil at offsets -2..2 (exact rational Vandermonde solve).
- madelung_sc_unit_cube: lim_{r->0} (psi(r) + 1/r) for the periodic
  potential psi of a unit mass in the unit cube with a uniform neutralizing
  background, in the attractive (psi ~ -1/r) convention.  Literature value of
  the simple-cubic Wigner/Madelung constant, 2.837297479480620 (e.g.
  Nijboer & De Wette 1957; Hummer, Pratt & Garcia 1995, xi_EW = -2.837297).
"""
import json
import os
from fractions import Fraction


def central_difference(offsets, deriv=1):
    n = len(offsets)
    A = [[Fraction(o) ** k for o in offsets] for k in range(n)]
    b = [Fraction(0)] * n
    b[deriv] = Fraction(1)  # d/dx x^1 at 0
    # Gaussian elimination over rationals
    for c in range(n):
        piv = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[piv] = A[piv], A[c]
        b[c], b[piv] = b[piv], b[c]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c] / A[c][c]
                A[r] = [a - f * p for a, p in zip(A[r], A[c])]
                b[r] -= f * b[c]
    return [float(b[i] / A[i][i]) for i in range(n)]


if __name__ == "__main__":
    out = {
        "central_difference_4th_order": central_difference([-2, -1, 0, 1, 2]),
        "madelung_sc_unit_cube": 2.837297479480620,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "analytic_values.json")
    json.dump(out, open(path, "w"), indent=1)
    print(out)
