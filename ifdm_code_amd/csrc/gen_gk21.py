"""Computes the 21-point Gauss-Kronrod rule on [-1, 1] in 60-digit precision
and prints it as the C++ tables of the qk21 rule in criteria.h.

Gauss nodes: roots of the Legendre polynomial P_10.  Kronrod nodes: roots of
the Stieltjes polynomial E_11, the monic odd polynomial of degree 11 with
int_{-1}^{1} E_11(x) P_10(x) x^k dx = 0 for k = 0..10.  Weights: the
interpolatory rule on all 21 nodes (exact for x^0..x^20; exactness to degree
31 then follows from the node choice and is checked in the tests)."""
import mpmath as mp

mp.mp.dps = 60


def legendre_coeffs(n):
    # coefficients of P_n, lowest degree first
    p0, p1 = [mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]
    if n == 0:
        return p0
    for k in range(1, n):
        a = [mp.mpf(0)] + [(2 * k + 1) * c for c in p1]
        b = p0 + [mp.mpf(0)] * (len(a) - len(p0))
        p0, p1 = p1, [(ai - k * bi) / (k + 1) for ai, bi in zip(a, b)]
    return p1


def moment(coeffs):  # int_{-1}^{1} sum c_i x^i dx
    return sum(c * (2 / mp.mpf(i + 1)) for i, c in enumerate(coeffs) if i % 2 == 0)


def polymul(a, b):
    r = [mp.mpf(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] += x * y
    return r


P10 = legendre_coeffs(10)
# E_11 = x^11 + c9 x^9 + c7 x^7 + c5 x^5 + c3 x^3 + c1 x ; conditions for odd k
unk = [1, 3, 5, 7, 9]
A = mp.matrix(5, 5)
bvec = mp.matrix(5, 1)
for row, k in enumerate([1, 3, 5, 7, 9]):
    xk = [mp.mpf(0)] * k + [mp.mpf(1)]
    base = polymul(P10, xk)
    for col, d in enumerate(unk):
        xd = [mp.mpf(0)] * d + [mp.mpf(1)]
        A[row, col] = moment(polymul(base, xd))
    x11 = [mp.mpf(0)] * 11 + [mp.mpf(1)]
    bvec[row] = -moment(polymul(base, x11))
cs = mp.lu_solve(A, bvec)
E11 = [mp.mpf(0)] * 12
E11[11] = mp.mpf(1)
for c, d in zip(cs, unk):
    E11[d] = c

gauss = sorted(mp.polyroots(list(reversed(P10)), maxsteps=200, extraprec=200), key=lambda z: -mp.re(z))
kron = sorted(mp.polyroots(list(reversed(E11)), maxsteps=200, extraprec=200), key=lambda z: -mp.re(z))
gauss = [mp.re(z) for z in gauss]
kron = [mp.re(z) for z in kron]
nodes = sorted(gauss + kron, key=lambda z: -z)          # 21 nodes, descending
pos = [x for x in nodes if x > mp.mpf("1e-40")] + [mp.mpf(0)]  # 10 positive + 0

# Kronrod weights: symmetric, 11 unknowns, exact for x^(2m), m = 0..10
M = mp.matrix(11, 11)
rhs = mp.matrix(11, 1)
for m in range(11):
    for i, x in enumerate(pos):
        M[m, i] = (2 * x ** (2 * m)) if i < 10 else (mp.mpf(1) if m == 0 else mp.mpf(0))
    rhs[m] = mp.mpf(2) / (2 * m + 1)
wk = mp.lu_solve(M, rhs)
# Gauss weights 2 / ((1 - x^2) P10'(x)^2)
dP = [i * c for i, c in enumerate(P10)][1:]
gpos = [x for x in sorted(gauss, key=lambda z: -z) if x > 0]
wg = [2 / ((1 - x * x) * sum(c * x ** i for i, c in enumerate(dP)) ** 2) for x in gpos]

def macro(name, vals):
    """a brace list of the values as one #define, three per line"""
    items = [mp.nstr(v, 25, min_fixed=-1, max_fixed=1) for v in vals]
    rows = [", ".join(items[i:i + 3]) for i in range(0, len(items), 3)]
    print(f"#define IFDM_GK21_{name} \\")
    print("    {" + ", \\\n     ".join(rows) + "}")


print("// 21-point Gauss-Kronrod rule on [-1,1], generated by gen_gk21.py.")
print("// Positive Kronrod nodes, descending, then 0 (odd indices are the Gauss nodes)")
macro("XGK", pos)
macro("WGK", wk)
print("// Gauss weights of the 10-point rule for the nodes xgk[1], xgk[3], .., xgk[9]")
macro("WG", wg)
