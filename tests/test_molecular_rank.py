"""The rank-3 factorisation the mixed form uses for the molecular operator of
Fourier modes m <= 2 (csrc/setup.hip: molecular_factors): the dense operator
w_i s(t_o, t_i) pair(q(t_o, t_i), o, i) of the reference's outer-product
channels (rt_sos_operator.c:102-119, with the coefficient selectors of
rt_sos_operator.c:383-406) equals E C G^T with the factors built from the
l = 2 basis at +o and the parities eps_p = eps_r = (-1)^m, eps_t = -(-1)^m,
to rounding, on the oracle's basis functions (Gauss nodes plus a zero-weight
node), and its rank is 3 for m = 0 and 2 for m = 1, 2."""
import ctypes as C

import numpy as np
import pytest

Q_SEL = [[0, 1, 8, 9, 4, 5], [1, 0, 9, 8, 5, 4], [6, 7, 10, 11, 12, 13],
         [7, 6, 11, 10, 13, 12], [2, 3, 14, 15, 16, 17], [3, 2, 15, 14, 17, 16]]
S_SEL = [[1, 1, 1, 1, -1, -1], [1, 1, 1, 1, 1, 1], [1, 1, 1, 1, -1, 1],
         [1, 1, 1, 1, -1, 1], [-1, 1, -1, -1, 1, 1], [-1, 1, 1, 1, 1, 1]]


def sel_type(col, n):
    t = col // n
    return 2 * t if t < 3 else 2 * (t - 3) + 1


@pytest.mark.parametrize("n", [8, 13])
@pytest.mark.parametrize("m", [0, 1, 2])
def test_molecular_operator_is_e_c_gt(oracle, n, m):
    D = C.POINTER(C.c_double)
    mu, w = np.zeros(n), np.zeros(n)
    oracle.oracle_gauss_legendre_pos(n, mu.ctypes.data_as(D), w.ctypes.data_as(D))
    order = np.argsort(mu)
    mu, w = mu[order], w[order]
    w[3] = 0.0   # a zero-weight (view) node
    dirs = 2 * n + 1
    mus = np.zeros(dirs)
    mus[n + 1:] = mu
    mus[:n] = -mu[::-1]
    plm, rrl, rtl = (np.zeros(3 * dirs) for _ in range(3))
    assert oracle.oracle_basis(2, m, dirs, mus.ctypes.data_as(D), plm.ctypes.data_as(D), rrl.ctypes.data_as(D),
                               rtl.ctypes.data_as(D)) == 0
    p2, r2, t2 = plm[2 * dirs:], rrl[2 * dirs:], rtl[2 * dirs:]
    b0, be2, ga2, al2 = (1.0 if m == 0 else 0.0), 0.4719, -0.6214, 2.8314

    def pair(q, o, i):
        po, pom, pi = p2[n + o], p2[n - o], p2[n + i]
        ro, rom, ri, rim = r2[n + o], r2[n - o], r2[n + i], r2[n - i]
        to, tom, ti, tim = t2[n + o], t2[n - o], t2[n + i], t2[n - i]
        return [b0 + be2 * pi * po, b0 + be2 * pi * pom, ga2 * pi * to, ga2 * pi * tom, ga2 * po * ti,
                ga2 * po * tim, ga2 * pi * ro, ga2 * pi * rom, ga2 * po * ri, ga2 * po * rim, al2 * ri * ro,
                al2 * ri * rom, al2 * ti * ro, al2 * ti * rom, al2 * to * ri, al2 * to * rim, al2 * ti * to,
                al2 * ti * tom][q]

    N = 6 * n
    A = np.zeros((N, N))
    for r in range(N):
        tr, o = sel_type(r, n), r - (r // n) * n + 1
        for c in range(N):
            tc, i = sel_type(c, n), c - (c // n) * n + 1
            if w[i - 1] != 0.0:
                A[r, c] = w[i - 1] * S_SEL[tr][tc] * pair(Q_SEL[tr][tc], o, i)
    ep = -1.0 if m & 1 else 1.0
    et = -ep
    E, G = np.zeros((N, 3)), np.zeros((N, 3))
    for c in range(N):
        t = sel_type(c, n)
        comp, h, o = t // 2, t % 2, c - (c // n) * n + 1
        if comp == 0:
            E[c] = [1.0, (ep if h else 1.0) * p2[n + o], 0.0]
        elif comp == 1:
            E[c] = [0.0, 0.0, (ep if h else 1.0) * r2[n + o]]
        else:
            E[c] = [0.0, 0.0, -(et if h else 1.0) * t2[n + o]]
        G[c] = w[o - 1] * E[c]
    Cm = np.array([[b0, 0.0, 0.0], [0.0, be2, ga2], [0.0, ga2, al2]])
    assert np.max(np.abs(A - E @ Cm @ G.T)) <= 1e-15 * np.max(np.abs(A))
    sv = np.linalg.svd(A, compute_uv=False)
    assert np.sum(sv > 1e-12 * sv[0]) == (3 if m == 0 else 2)
