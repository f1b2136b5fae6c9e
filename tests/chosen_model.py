"""numpy FP64 model of the product's FP32 formulation of the repulsive term (test infrastructure).

The reference interpolates charges [1, x, y, x^2+y^2] (tsne.cpp:1057-1062) and recombines
(x_i phi_1 - phi_2) and (1+|y_i|^2) phi_1 - 2 y_i.phi_23 + phi_4 (tsne.cpp:1101-1110, :1148-1155).
Those sums cancel: at |y| ~ 80 each term is 10^3-10^4 times its result, which FP32 cannot carry.

The product evaluates the SAME interpolated quantities without cancellation by writing every
point-node difference relative to the node:  y_i - y_j = e_in + (yt_n - yt_n') - e_jn'  with
e_in = y_i - yt_n.  With grid sums A_n = sum_i L_i(n), b_n = sum_i L_i(n) e_in, c_n = sum_i L_i(n)|e_in|^2:

  sum_j Kt(i,j)(y_i - y_j) = sum_n L_i(n) [ e_in (K*A)_n + (K_D*A)_n - (K*b)_n ]        (forces)
  Z + N = <A, K1*A> + 2<c, K*A> + 4<b, K_D*A> - 2<b, K*b>                             (df = 1)
  Z + N = <A, Kg*A>                                                                    (df != 1)

where K_D(d) = K(d) * d * h is the odd kernel and K1 = K (1+|d h|^2) = (1+r^2)^-1 on the node lattice.
This module checks that identity against the oracle in FP64 before it is trusted in FP32.
"""
import numpy as np

import orc


def _stages(m):
    """Stages of the product's FFT plan for m = 2^a 3^b (radices 16, 8, 4, 2 then 9, 3)."""
    s = 0
    while m % 16 == 0:
        s, m = s + 1, m // 16
    for r in (8, 4, 2):
        if m % r == 0:
            s, m = s + 1, m // r
    while m % 9 == 0:
        s, m = s + 1, m // 9
    while m % 3 == 0:
        s, m = s + 1, m // 3
    return s


def fft_size(n):
    """The product's FFT length for a linear convolution of n1d nodes (n = 2 n1d - 1): the smallest
    2^a 3^b >= n, or for lines up to 512 the 2^a 3^b in [n, next power of two] with the least
    length x stages.  Any length >= 2 n1d - 1 gives the same convolution as the reference's 2 n1d."""
    pw = 1
    while pw < n:
        pw *= 2
    best = None
    p3 = 1
    while p3 <= pw:
        m = p3
        while m < n:
            m *= 2
        if m <= pw:
            cost = (m * _stages(m), m) if pw <= 512 else (m, m)
            if best is None or cost < best[0]:
                best = (cost, m)
        p3 *= 3
    return best[1]


def _lagrange(t, sp):
    p = len(sp)
    den = np.ones(p)
    for i in range(p):
        for j in range(p):
            if i != j:
                den[i] *= sp[i] - sp[j]
    L = np.ones((p, t.size))
    for j in range(p):
        for k in range(p):
            if j != k:
                L[j] *= t - sp[k]
        L[j] /= den[j]
    return L


def _spacings(p):
    h = 1 / p
    sp = np.empty(p)
    sp[0] = h / 2
    for i in range(1, p):
        sp[i] = sp[i - 1] + h
    return sp


def chosen_gradient(P, Y, df=1.0, p=3, ipi=1.0, min_num_intervals=50, return_parts=False):
    Y = np.asarray(Y, dtype=np.float64)
    D = 1 if Y.ndim == 1 or Y.shape[1] == 1 else 2
    Y2 = Y.reshape(-1, D)
    N = Y2.shape[0]
    if D == 2:
        mn, mx = orc.minmax_quirk(Y2)
        nb = orc.oracle().orc_n_boxes_2d(mn, mx, ipi, min_num_intervals)
    else:
        mn, mx = Y2.min(), Y2.max()
        nb = orc.oracle().orc_n_boxes_1d(mn, mx, ipi, min_num_intervals)
    bw0 = (mx - mn) / nb
    bw = ((1 * bw0 + mn) - (0 * bw0 + mn))
    n1d = nb * p
    Mp = fft_size(2 * n1d - 1)
    h = (1 / p) * bw0
    sp = _spacings(p)

    idx, L, E = [], [], []
    for d in range(D):
        x = Y2[:, d]
        bi = ((x - mn) / bw).astype(np.int64)
        bi = np.minimum(bi, nb - 1)
        if D == 2:
            bi = np.maximum(bi, 0)
        t = (x - (bi * bw0 + mn)) / bw
        idx.append(bi)
        L.append(_lagrange(t, sp))
        E.append((t[None, :] - sp[:, None]) * bw0)  # e = x - node, [p, N]

    # kernels on lattice offsets, embedded with wrap-around in Mp
    off = np.arange(-(n1d - 1), n1d)
    kq_kind = orc.SQUARED_CAUCHY if df == 1.0 else orc.SQUARED_GENERAL
    if D == 2:
        du, dv = np.meshgrid(off * h, off * h, indexing="ij")
        r2 = du * du + dv * dv
    else:
        du = off * h
        r2 = du * du
    if df == 1.0:
        Kq = (1.0 + r2) ** -2.0
        K1 = (1.0 + r2) ** -1.0
    else:
        Kq = (1.0 + r2 / df) ** -(df + 1.0)
        K1 = (1.0 + r2 / df) ** -df

    def embed(k):
        out = np.zeros((Mp,) * D)
        ii = off % Mp
        if D == 2:
            out[np.ix_(ii, ii)] = k
        else:
            out[ii] = k
        return np.fft.fftn(out)

    Kq_h, K1_h = embed(Kq), embed(K1)
    KD_h = [embed(Kq * (du if D == 1 else (du if d == 0 else dv))) for d in range(D)]

    # spread
    shape = (n1d,) * D
    A = np.zeros(shape)
    B = [np.zeros(shape) for _ in range(D)]
    Cg = np.zeros(shape)
    if D == 2:
        for a in range(p):
            for b in range(p):
                w = L[0][a] * L[1][b]
                ij = (idx[0] * p + a, idx[1] * p + b)
                np.add.at(A, ij, w)
                np.add.at(B[0], ij, w * E[0][a])
                np.add.at(B[1], ij, w * E[1][b])
                np.add.at(Cg, ij, w * (E[0][a] ** 2 + E[1][b] ** 2))
    else:
        for a in range(p):
            w = L[0][a]
            ij = idx[0] * p + a
            np.add.at(A, ij, w)
            np.add.at(B[0], ij, w * E[0][a])
            np.add.at(Cg, ij, w * E[0][a] ** 2)

    def fwd(g):
        out = np.zeros((Mp,) * D)
        out[tuple(slice(0, n1d) for _ in range(D))] = g
        return np.fft.fftn(out)

    def inv(gh):
        return np.real(np.fft.ifftn(gh))[tuple(slice(0, n1d) for _ in range(D))]

    Ah, Ch = fwd(A), fwd(Cg)
    Bh = [fwd(b) for b in B]
    V1 = inv(Kq_h * Ah)
    G = [inv(KD_h[d] * Ah - Kq_h * Bh[d]) for d in range(D)]

    def ip(fh, kh, gh):  # <f, k*g> by Parseval
        return np.real(np.sum(np.conj(fh) * kh * gh)) / Mp ** D

    if df == 1.0:
        Zf = ip(Ah, K1_h, Ah) + 2 * ip(Ch, Kq_h, Ah)
        for d in range(D):
            Zf += 4 * ip(Bh[d], KD_h[d], Ah) - 2 * ip(Bh[d], Kq_h, Bh[d])
    else:
        Zf = ip(Ah, K1_h, Ah)
    Z = Zf - N

    neg = np.zeros((N, D))
    if D == 2:
        for a in range(p):
            for b in range(p):
                w = L[0][a] * L[1][b]
                u, v = idx[0] * p + a, idx[1] * p + b
                neg[:, 0] += w * (E[0][a] * V1[u, v] + G[0][u, v])
                neg[:, 1] += w * (E[1][b] * V1[u, v] + G[1][u, v])
    else:
        for a in range(p):
            u = idx[0] * p + a
            neg[:, 0] += L[0][a] * (E[0][a] * V1[u] + G[0][u])
    pos = orc.attractive(P, Y2, df)
    dC = pos - neg / Z
    if return_parts:
        return dC, Z, dict(nb=nb, Mp=Mp, n1d=n1d)
    return dC, Z
