"""CPU oracle for the TwoPunctures initial data (test infrastructure only:
imported by tests/, never by the product path).

A numpy restatement of the two-puncture problem of
BSSN_GR/src/TwoPunctures.cpp (driver, :24-498) and its solver
(TPEquations.cpp, TPCoordTransf.cpp, TPFuncAndJacobian.cpp, TPNewton.cpp):

  psi = 1 + m+/(2 r+) + m-/(2 r-) + u,   Lap u + (1/8) A_ij A^ij psi^-7 = 0,

with the Bowen-York A_ij of TPEquations.cpp:81-145, discretised on the
collocation grid of TPFuncAndJacobian.cpp:221-225 (Chebyshev-Gauss in A and
B, equispaced phi) in the prolate spheroidal coordinates of
TPCoordTransf.cpp:18-50, and solved by Newton's method with a dense numpy
solve (small resolutions only).  The reference cannot be built here (it
needs GSL and MPI headers), so parity is pinned by what the mathematics
fixes, in the tests: the Brill-Lindquist limit (u = 0 without momenta and
spins), the ADM linear and angular momentum surface integrals of A_ij
(= the P and S put in), and the finite-difference Hamiltonian constraint of
the interpolated solution vanishing at the solver's spectral accuracy.
"""
import numpy as np


def cheb_gauss(n):
    """Chebyshev-Gauss nodes -cos(pi (2i+1) / (2n)) (TPFuncAndJacobian.cpp:221-224)."""
    return -np.cos(np.pi * (2 * np.arange(n) + 1) / (2 * n))


def diff_matrix(x):
    """Differentiation matrix of the polynomial interpolant at nodes x."""
    n = len(x)
    w = np.array([1.0 / np.prod([x[j] - x[k] for k in range(n) if k != j]) for j in range(n)])
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = (w[j] / w[i]) / (x[i] - x[j])
        D[i, i] = -D[i].sum()
    return D


def fourier_d2(n):
    """Second derivative of the trigonometric interpolant at 2 pi k / n."""
    if n < 2:
        return np.zeros((n, n))
    h = 2 * np.pi / n
    D = np.zeros((n, n))
    for k in range(n):
        for l in range(n):
            d = k - l
            if d == 0:
                D[k, l] = -np.pi ** 2 / (3 * h * h) - 1.0 / 6 if n % 2 == 0 else -(n * n - 1) / 12.0
            elif n % 2 == 0:
                D[k, l] = -((-1) ** d) / (2 * np.sin(d * h / 2) ** 2)
            else:
                D[k, l] = -((-1) ** d) * np.cos(d * h / 2) / (2 * np.sin(d * h / 2) ** 2)
    return D


def by_aij(p, x, y, z, regular=False):
    """Bowen-York A_ij (3, 3, ...) of both punctures at coordinates relative
    to the centre (TPEquations.cpp:81-145; regular: the grid's
    sqrt(r^4 + eps^4) distance regularisation, TPEquations.cpp:87-94)."""
    x, y, z = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float))
    A = np.zeros((3, 3) + x.shape)
    b = p["par_b"]
    for c, P, S in ((b, p["par_P_plus"], p["par_S_plus"]), (-b, p["par_P_minus"], p["par_S_minus"])):
        d = [x - c, y, z]
        r2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
        if regular:
            r2 = np.sqrt(r2 ** 2 + p["TP_epsilon"] ** 4)
            r2 = np.maximum(r2, p["TP_Tiny"] ** 2)
        r = np.sqrt(r2)
        n = [di / r for di in d]
        nP = sum(n[i] * P[i] for i in range(3))
        nS = [n[1] * S[2] - n[2] * S[1], n[2] * S[0] - n[0] * S[2], n[0] * S[1] - n[1] * S[0]]
        for i in range(3):
            for j in range(3):
                a = 1.5 * (P[i] * n[j] + P[j] * n[i] + nP * n[i] * n[j]) / r2 - 3.0 * (nS[i] * n[j] + nS[j] * n[i]) / (r * r2)
                if i == j:
                    a = a - 1.5 * nP / r2
                A[i, j] += a
    return A


def to_abphi(b, x, y, z):
    """(A, B, phi) of points relative to the centre: cosh X = (r+ + r-)/(2b),
    cos R = (r- - r+)/(2b), A = tanh(X/2), B = tan((R - pi/2)/2)."""
    rp = np.sqrt((x - b) ** 2 + y ** 2 + z ** 2)
    rm = np.sqrt((x + b) ** 2 + y ** 2 + z ** 2)
    s = np.maximum((rp + rm) / (2 * b), 1.0)
    c = np.clip((rm - rp) / (2 * b), -1.0, 1.0)
    A = np.sqrt((s - 1) / (s + 1))
    sm, sp = np.sqrt(1 - c), np.sqrt(1 + c)
    return A, (sm - sp) / (sm + sp), np.arctan2(z, y)


class TPOracle:
    """The collocation problem and its dense Newton solve."""

    def __init__(self, p, mp=None, mm=None):
        self.p = p
        n1, n2, n3 = p["npoints_A"], p["npoints_B"], p["npoints_phi"]
        self.n = (n1, n2, n3)
        b = p["par_b"]
        self.mp = p["par_m_plus"] if mp is None else mp
        self.mm = p["par_m_minus"] if mm is None else mm
        a, B = cheb_gauss(n1), cheb_gauss(n2)
        ph = 2 * np.pi * np.arange(n3) / n3
        self.a, self.B, self.ph = a, B, ph
        DA = 2 * diff_matrix(a)
        DB = diff_matrix(B)
        D2p = fourier_d2(n3)
        I1, I2, I3 = np.eye(n1), np.eye(n2), np.eye(n3)

        def kron(P3, P2, P1):  # index i + n1 (j + n2 k)
            return np.kron(P3, np.kron(P2, P1))
        MA, MAA = kron(I3, I2, DA), kron(I3, I2, DA @ DA)
        MB, MBB = kron(I3, DB, I1), kron(I3, DB @ DB, I1)
        Mpp = kron(D2p, I2, I1)
        At = np.tile((a + 1) / 2, n2 * n3)
        Bv = np.tile(np.repeat(B, n1), n3)
        phv = np.repeat(ph, n1 * n2)
        self.At = At
        AX, BR = (1 - At ** 2) / 2, (1 + Bv ** 2) / 2
        AXX, BRR = -At * AX, Bv * BR
        coth, sh = (1 + At ** 2) / (2 * At), 2 * At / (1 - At ** 2)
        cot, sn = -2 * Bv / (1 - Bv ** 2), (1 - Bv ** 2) / (1 + Bv ** 2)
        den = b * b * (sh ** 2 + sn ** 2)
        T = np.diag(At - 1)
        I = np.eye(len(At))
        d = np.diag
        # flat Laplacian of u = (A - 1) V in prolate spheroidal coordinates
        self.L = (d(AX ** 2 / den) @ (2 * MA + T @ MAA) + d((AXX + coth * AX) / den) @ (I + T @ MA)
                  + d(BR ** 2 / den) @ (T @ MBB) + d((BRR + cot * BR) / den) @ (T @ MB)
                  + d(1 / (b * b * sh ** 2 * sn ** 2)) @ (T @ Mpp))
        chX, cR = (1 + At ** 2) / (1 - At ** 2), -2 * Bv / (1 + Bv ** 2)
        x, rho = b * chX * cR, b * sh * sn
        self.xyz = (x, rho * np.cos(phv), rho * np.sin(phv))
        Aij = by_aij(p, *self.xyz)
        self.AA = (Aij ** 2).sum(axis=(0, 1))
        self.V = np.zeros(len(At))
        self.set_masses(self.mp, self.mm)

    def set_masses(self, mp, mm):
        b = self.p["par_b"]
        x, y, z = self.xyz
        self.mp, self.mm = mp, mm
        self.psiBL = 1 + 0.5 * mp / np.sqrt((x - b) ** 2 + y ** 2 + z ** 2) + 0.5 * mm / np.sqrt((x + b) ** 2 + y ** 2 + z ** 2)

    def newton(self, tol, maxit):
        """TPNewton.cpp's iteration with a dense solve; stops when the
        largest update is below tol."""
        for _ in range(maxit):
            psi = self.psiBL + (self.At - 1) * self.V
            F = self.L @ self.V + 0.125 * self.AA / psi ** 7
            J = self.L + np.diag(-0.875 * self.AA / psi ** 8 * (self.At - 1))
            dV = np.linalg.solve(J, F)
            self.V -= dV
            if np.abs(dV).max() < tol:
                break

    def eval_V(self, A, B, phi):
        """V at (A, B, phi) by Chebyshev x Chebyshev x trigonometric
        interpolation of the collocation values."""
        n1, n2, n3 = self.n
        Vg = self.V.reshape(n3, n2, n1)
        Ta = np.polynomial.chebyshev.chebvander(self.a, n1 - 1)
        Tb = np.polynomial.chebyshev.chebvander(self.B, n2 - 1)
        ca = np.linalg.solve(Ta, Vg.transpose(2, 1, 0).reshape(n1, -1)).reshape(n1, n2, n3)
        cb = np.linalg.solve(Tb, ca.transpose(1, 0, 2).reshape(n2, -1)).reshape(n2, n1, n3)
        A, B, phi = np.atleast_1d(A), np.atleast_1d(B), np.atleast_1d(phi)
        ta = np.polynomial.chebyshev.chebvander(2 * A - 1, n1 - 1)
        tb = np.polynomial.chebyshev.chebvander(B, n2 - 1)
        vk = np.einsum('pl,pm,lmk->pk', tb, ta, cb)
        # trigonometric interpolation in phi (real DFT)
        out = vk.mean(axis=1)
        M = (n3 - 1) // 2
        for m in range(1, M + 1):
            cm = 2 * (vk * np.cos(m * self.ph)).mean(axis=1)
            sm = 2 * (vk * np.sin(m * self.ph)).mean(axis=1)
            out = out + cm * np.cos(m * phi) + sm * np.sin(m * phi)
        if n3 % 2 == 0 and n3 >= 2:
            out = out + (vk * np.cos(n3 // 2 * self.ph)).mean(axis=1) * np.cos(n3 // 2 * phi)
        return out

    def u(self, x, y, z):
        """u at points relative to the centre."""
        A, B, ph = to_abphi(self.p["par_b"], np.asarray(x, float), np.asarray(y, float), np.asarray(z, float))
        return (A - 1) * self.eval_V(A, B, ph)


def bssn_point(p, mp, mm, x, y, z, u):
    """The 24 BSSN variables (dict of arrays) from u at points relative to
    the centre (TwoPunctures.cpp:270-498 with adm2bssn): psi with the
    regularised distances, lapse of initial_lapse, gd = psi^4 delta,
    K_ij = A_ij / psi^2, chi = det(gd)^(-1/3), gt = chi gd, K = trace,
    At = chi (K_ij - gd_ij K / 3)."""
    b, eps, tiny, R = p["par_b"], p["TP_epsilon"], p["TP_Tiny"], p["TP_Extend_Radius"]

    def rr(c):
        r = np.sqrt((x - c) ** 2 + y ** 2 + z ** 2)
        return np.maximum((r ** 4 + eps ** 4) ** 0.25, tiny)

    def inv(r):
        if R > 0:
            ext = 3.0 / 8 * r ** 4 / R ** 5 - 5.0 / 4 * r ** 2 / R ** 3 + 15.0 / 8 / R
            return np.where(r < R, ext, 1.0 / r)
        return 1.0 / r
    ip, im = inv(rr(b)), inv(rr(-b))
    psi_bl = 1 + 0.5 * mp * ip + 0.5 * mm * im
    psi1 = psi_bl + u
    lapse = p["initial_lapse"]
    if lapse in (0, 1):
        alpha = (1 - 0.5 * mp * ip - 0.5 * mm * im) / (1 + 0.5 * mp * ip + 0.5 * mm * im)
        if lapse == 1:
            alpha = 0.5 * (1 + alpha)
    elif lapse in (2, 3):
        alpha = psi_bl ** p["initial_lapse_psi_exponent"]
        if lapse == 3:
            alpha = 2 / (1 + alpha)
    else:
        alpha = np.ones_like(psi1)
    Aij = by_aij(p, x, y, z, regular=True)
    p2 = psi1 * psi1
    p4 = p2 * p2
    chi = (1.0 / (p4 * p4 * p4)) ** (1.0 / 3.0)
    trK = (Aij[0, 0] + Aij[1, 1] + Aij[2, 2]) / p4 / p2
    out = {"alpha": alpha, "chi": chi, "K": trK}
    sym = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for q, (r, c) in enumerate(sym):
        gd = p4 if r == c else 0.0 * p4
        out["gt%d" % q] = chi * gd
        out["At%d" % q] = chi * (Aij[r, c] / p2 - gd * trK / 3.0)
    return out


def adm_momenta(p, radius=40.0, nth=64, nph=128):
    """ADM linear and angular momentum of the Bowen-York curvature by
    surface integrals over a sphere of the given radius around the centre:
    P_i = (1/8pi) oint A_ij n^j dS, J_i = (1/8pi) oint eps_ijk x^j A^kl n_l dS
    (Gauss-Legendre in cos(theta), trapezoid in phi)."""
    mu, w = np.polynomial.legendre.leggauss(nth)
    phi = 2 * np.pi * np.arange(nph) / nph
    MU, PH = np.meshgrid(mu, phi, indexing="ij")
    st = np.sqrt(1 - MU ** 2)
    n = np.stack([st * np.cos(PH), st * np.sin(PH), MU])
    X = radius * n
    A = by_aij(p, X[0], X[1], X[2])
    W = np.outer(w, np.full(nph, 2 * np.pi / nph)) * radius ** 2
    An = np.einsum('ij...,j...->i...', A, n)
    P = np.array([(An[i] * W).sum() for i in range(3)]) / (8 * np.pi)
    xA = np.cross(X, An, axis=0)
    J = np.array([(xA[i] * W).sum() for i in range(3)]) / (8 * np.pi)
    return P, J
