"""Independent numpy formulation of the LDG right-hand side (test helper).

This is a second, independently written statement of the same discrete
operator, used to pin the C oracle.  It shares no code or data with the
oracle: the 1D basis comes from numpy polynomials (Lagrange products, numpy
Gauss-Legendre), the weak form is integrated by tensor-product quadrature
instead of assembled dense matrices, the Roe flux is built from its
eigen-decomposition with a numerical 5x5 solve (not the compact formula),
and the viscous flux goes through the pressure gradient rather than the
internal-energy gradient.
"""
import numpy as np
from numpy.polynomial import legendre as npleg

GAMMA = 1.4


def cgl_nodes(p):
    """Chebyshev-Gauss-Lobatto nodes ascending on [-1, 1] (closed form)."""
    return -np.cos(np.pi * np.arange(p + 1) / p)


def lagrange_ops(p):
    """1D Lagrange basis on the CGL nodes evaluated (I) and differentiated
    (D) at the (p+1)-point Gauss rule; also returns the Gauss rule and the
    basis evaluated at -1 and +1."""
    r = cgl_nodes(p)
    g, w = npleg.leggauss(p + 1)
    n = p + 1
    I = np.zeros((n, n))
    D = np.zeros((n, n))
    ends = np.zeros((2, n))
    for i in range(n):
        others = np.delete(r, i)
        poly = np.polynomial.Polynomial.fromroots(others)
        poly = poly / poly(r[i])
        I[:, i] = poly(g)
        D[:, i] = poly.deriv()(g)
        ends[:, i] = poly(np.array([-1.0, 1.0]))
    return r, g, w, I, D, ends


def euler_flux(U, gamma=GAMMA):
    """Euler flux, U (..., 5) -> F (..., 3, 5) with F[..., l, s]."""
    rho = U[..., 0]
    v = U[..., 1:4] / rho[..., None]
    E = U[..., 4]
    pr = (gamma - 1.0) * (E - 0.5 * rho * np.sum(v * v, axis=-1))
    F = np.zeros(U.shape[:-1] + (3, 5))
    for l in range(3):
        F[..., l, 0] = rho * v[..., l]
        for i in range(3):
            F[..., l, 1 + i] = rho * v[..., i] * v[..., l] + (pr if i == l else 0.0)
        F[..., l, 4] = (E + pr) * v[..., l]
    return F


def viscous_flux(U, G, Re, Pr, gamma=GAMMA):
    """Navier-Stokes viscous flux via primitive variables.

    U (..., 5); G (..., 3, 5) with G[..., l, s] = d U_s / d x_l.
    Returns Fv (..., 3, 5).  Heat flux kappa grad(T) written through the
    pressure: e = p/((gamma-1) rho), kappa = mu gamma / Pr on e.
    """
    rho = U[..., 0]
    m = U[..., 1:4]
    E = U[..., 4]
    v = m / rho[..., None]
    mu = 1.0 / Re
    # grad v: dv[..., i, l] = d v_i / d x_l
    dv = (np.swapaxes(G[..., :, 1:4], -1, -2) - v[..., :, None] * G[..., None, :, 0]) / rho[..., None, None]
    div = dv[..., 0, 0] + dv[..., 1, 1] + dv[..., 2, 2]
    tau = mu * (dv + np.swapaxes(dv, -1, -2))
    for i in range(3):
        tau[..., i, i] -= mu * 2.0 / 3.0 * div
    pr = (gamma - 1.0) * (E - 0.5 * np.sum(m * m, axis=-1) / rho)
    # grad p = (gamma-1)(grad E - v.grad m + |v|^2/2 grad rho)
    gp = (gamma - 1.0) * (G[..., :, 4] - np.einsum("...i,...li->...l", v, G[..., :, 1:4])
                          + 0.5 * np.sum(v * v, axis=-1)[..., None] * G[..., :, 0])
    ge = (gp / rho[..., None] - pr[..., None] * G[..., :, 0] / rho[..., None] ** 2) / (gamma - 1.0)
    F = np.zeros(U.shape[:-1] + (3, 5))
    for l in range(3):
        F[..., l, 1:4] = tau[..., :, l]
        F[..., l, 4] = np.einsum("...i,...i->...", v, tau[..., :, l]) + mu * gamma / Pr * ge[..., l]
    return F


def roe_flux_eig(uN, uK, n, gamma=GAMMA):
    """Roe flux for side K with outward normal n, from the eigen-decomposition
    of the Roe-averaged Jacobian: F = (F_K + F_N).n/2 - sum_i |lam_i| a_i r_i / 2
    with R a = uN - uK solved numerically.  No entropy fix."""
    uN = np.asarray(uN, float)
    uK = np.asarray(uK, float)
    n = np.asarray(n, float)
    FN = euler_flux(uN, gamma)
    FK = euler_flux(uK, gamma)
    avg = 0.5 * (np.einsum("l,ls->s", n, FN) + np.einsum("l,ls->s", n, FK))
    sN, sK = np.sqrt(uN[0]), np.sqrt(uK[0])

    def prim(u):
        rho = u[0]
        v = u[1:4] / rho
        pr = (gamma - 1.0) * (u[4] - 0.5 * rho * v @ v)
        return v, (u[4] + pr) / rho

    vN, HN = prim(uN)
    vK, HK = prim(uK)
    vt = (sN * vN + sK * vK) / (sN + sK)
    Ht = (sN * HN + sK * HK) / (sN + sK)
    c = np.sqrt((gamma - 1.0) * (Ht - 0.5 * vt @ vt))
    qn = vt @ n
    # tangents
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    t1 = np.cross(n, a)
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    R = np.zeros((5, 5))
    R[:, 0] = np.concatenate([[1.0], vt - c * n, [Ht - c * qn]])
    R[:, 1] = np.concatenate([[1.0], vt, [0.5 * vt @ vt]])
    R[:, 2] = np.concatenate([[0.0], t1, [vt @ t1]])
    R[:, 3] = np.concatenate([[0.0], t2, [vt @ t2]])
    R[:, 4] = np.concatenate([[1.0], vt + c * n, [Ht + c * qn]])
    lam = np.abs(np.array([qn - c, qn, qn, qn, qn + c]))
    alpha = np.linalg.solve(R, uN - uK)
    return avg - 0.5 * R @ (lam * alpha)


class NumpyLDG:
    """Weak-form LDG right-hand side on an nx*ny*nz periodic box, order p.

    State layout matches the reference: u[k, s, iN] with
    iN = ix + n iy + n^2 iz and k = ex + nx ey + nx ny ez.
    """

    def __init__(self, p, nx, ny=None, nz=None, xmin=None, xmax=None, Re=1600.0, Pr=0.71, gamma=GAMMA):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self.p, self.n = p, p + 1
        self.ne = (nx, ny, nz)
        xmin = np.full(3, -np.pi) if xmin is None else np.asarray(xmin, float)
        xmax = np.full(3, np.pi) if xmax is None else np.asarray(xmax, float)
        self.dx = (xmax - xmin) / np.array(self.ne)
        self.alpha = self.dx / 2
        self.J = np.prod(self.alpha)
        self.Re, self.Pr, self.gamma = Re, Pr, gamma
        self.r, self.g, self.w, self.I, self.D, self.ends = lagrange_ops(p)
        self.M1 = self.I.T @ np.diag(self.w) @ self.I
        self.M1inv = np.linalg.inv(self.M1)
        self.K = nx * ny * nz

    # layout helpers: (K,5,dofs) <-> (ez,ey,ex,5,iz,iy,ix)
    def to_grid(self, u):
        nx, ny, nz = self.ne
        n = self.n
        return u.reshape(nz, ny, nx, 5, n, n, n)

    def from_grid(self, G):
        return G.reshape(self.K, 5, self.n ** 3)

    def _interp3(self, G):
        I = self.I
        return np.einsum("ai,bj,ck,...ijk->...abc", I, I, I, G, optimize=True)

    def _face_vals(self, G, l, side):
        """Nodal trace on face (direction l, side 0 = minus, 1 = plus):
        returns array (..., a2, a1) over the two remaining node indices in
        (iz,iy,ix) order with the normal axis removed."""
        idx = 0 if side == 0 else self.n - 1
        ax = -1 - l   # ix -> -1, iy -> -2, iz -> -3
        return np.take(G, idx, axis=ax)

    def _interp2(self, F):
        I = self.I
        return np.einsum("ai,bj,...ij->...ab", I, I, F, optimize=True)

    def _neighbor(self, F, l, side):
        """Values of the neighbour across face (l, side): roll the element
        grid.  F has element axes (ez, ey, ex) first."""
        shift = 1 if side == 0 else -1
        return np.roll(F, shift, axis=2 - l)

    def _minv(self, R):
        Mi = self.M1inv
        return np.einsum("ai,bj,ck,...ijk->...abc", Mi, Mi, Mi, R, optimize=True) / self.J

    def _vol_test(self, F, l):
        """sum_q w_q J/alpha_l d_l phi_i(q) F(q); F (..., qz,qy,qx)."""
        I, D, w = self.I, self.D, self.w
        W3 = np.einsum("a,b,c->abc", w, w, w)
        ops = [I, I, I]
        ops[2 - l] = D
        return self.J / self.alpha[l] * np.einsum("ai,bj,ck,...abc->...ijk", ops[0], ops[1], ops[2], F * W3, optimize=True)

    def _face_test(self, Fr, l, side):
        """Lift of face-quadrature values Fr (..., qa2, qa1) onto the nodes:
        returns (..., iz, iy, ix) array nonzero on the face only."""
        I, w = self.I, self.w
        W2 = np.outer(w, w)
        lifted = self.J / self.alpha[l] * np.einsum("ai,bj,...ab->...ij", I, I, Fr * W2, optimize=True)
        out_shape = lifted.shape[:-2] + (self.n, self.n, self.n)
        out = np.zeros(out_shape)
        # face nodal basis: ell_i(+-1) = ends[side, i]
        e = self.ends[side]
        if l == 0:
            out += lifted[..., :, :, None] * e
        elif l == 1:
            out += lifted[..., :, None, :] * e[:, None]
        else:
            out += lifted[..., None, :, :] * e[:, None, None]
        return out

    def gradient(self, u):
        """LDG auxiliary gradient Q[l] (3, K, 5, dofs)."""
        G = self.to_grid(u)
        Gq = self._interp3(G)
        Q = []
        for l in range(3):
            R = -self._vol_test(Gq, l)
            for side in (0, 1):
                nl = -1.0 if side == 0 else 1.0
                own = self._interp2(self._face_vals(G, l, side))
                nb = self._interp2(self._neighbor(self._face_vals(G, l, 1 - side), l, side))
                uhat = own if nl > 0 else nb
                R = R + self._face_test(uhat * nl, l, side)
            Q.append(self.from_grid(self._minv(R)))
        return np.stack(Q)

    def rhs(self, u, return_grad=False):
        Q = self.gradient(u)
        G = self.to_grid(u)
        GQ = [self.to_grid(Q[l]) for l in range(3)]
        Uq = np.moveaxis(self._interp3(G), 3, -1)              # (...,qz,qy,qx,5)
        Qq = np.stack([np.moveaxis(self._interp3(GQ[l]), 3, -1) for l in range(3)], axis=-2)  # (...,3,5)
        Fc = euler_flux(Uq, self.gamma)
        Fv = viscous_flux(Uq, Qq, self.Re, self.Pr, self.gamma)
        Fvol = np.moveaxis(Fc - Fv, -1, 3)                        # (ez,ey,ex,5,qz,qy,qx,3)
        R = sum(self._vol_test(Fvol[..., l], l) for l in range(3))
        for l in range(3):
            for side in (0, 1):
                nvec = np.zeros(3)
                nvec[l] = -1.0 if side == 0 else 1.0
                uK = np.moveaxis(self._interp2(self._face_vals(G, l, side)), 3, -1)
                uN = np.moveaxis(self._interp2(self._neighbor(self._face_vals(G, l, 1 - side), l, side)), 3, -1)
                if nvec.sum() > 0:
                    qs = [self._neighbor(self._face_vals(GQ[m], l, 1 - side), l, side) for m in range(3)]
                else:
                    qs = [self._face_vals(GQ[m], l, side) for m in range(3)]
                Qt = np.stack([np.moveaxis(self._interp2(q), 3, -1) for q in qs], axis=-2)
                shp = uK.shape[:-1]
                uKf = uK.reshape(-1, 5)
                uNf = uN.reshape(-1, 5)
                roe = np.array([roe_flux_eig(a, b, nvec, self.gamma) for a, b in zip(uNf, uKf)]).reshape(shp + (5,))
                Fvf = viscous_flux(uK, Qt, self.Re, self.Pr, self.gamma)
                fv_star = -np.einsum("...ls,l->...s", Fvf, nvec)
                Fstar = np.moveaxis(roe + fv_star, -1, 3)
                R = R - self._face_test(Fstar, l, side)
        k = self.from_grid(self._minv(R))
        return (k, Q) if return_grad else k
