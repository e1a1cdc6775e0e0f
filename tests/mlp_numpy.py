"""Independent numpy restatement of the closure network and of the L2
projection between orders, used as the checker of libldg.so's closure
kernels.  Test infrastructure only.

Network: H1 = silu(X W1^T + b1), H2 = silu(H1 W2^T + b2), Y = H2 W3^T + b3,
loss = mean((Y - T)^2).  With bf16=True every matrix-product operand is
rounded to BF16 (round to nearest even) as the GPU kernels do before their
FP32-accumulating matrix-core products; accumulation here is FP64.
"""
import numpy as np


def bf16(a):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(params, nin, hidden, nout):
    sizes = [hidden * nin, hidden, hidden * hidden, hidden, nout * hidden, nout]
    shapes = [(hidden, nin), (hidden,), (hidden, hidden), (hidden,), (nout, hidden), (nout,)]
    out, off = [], 0
    for n, s in zip(sizes, shapes):
        out.append(np.asarray(params[off:off + n], dtype=np.float64).reshape(s))
        off += n
    return out


def _f32(a, on):
    return np.asarray(a, dtype=np.float32).astype(np.float64) if on else a


def forward(params, X, nin, hidden, nout, bf=False):
    W1, b1, W2, b2, W3, b3 = unpack(params, nin, hidden, nout)
    Z1 = _r(X, bf) @ _r(W1, bf).T + b1
    H1 = _f32(silu(Z1), bf)          # the GPU stores activations in FP32
    Z2 = _r(H1, bf) @ _r(W2, bf).T + b2
    H2 = _f32(silu(Z2), bf)
    Y = _r(H2, bf) @ _r(W3, bf).T + b3
    return Y, (Z1, H1, Z2, H2)


def loss(params, X, T, nin, hidden, nout):
    Y, _ = forward(params, X, nin, hidden, nout)
    return float(np.mean((Y - T) ** 2))


def gradient(params, X, T, nin, hidden, nout, bf=False):
    """Gradient of mean((Y - T)^2) by the chain rule, in the same order as
    the GPU backward pass (each product's operands rounded when bf)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, nin, hidden, nout)
    Y, (Z1, H1, Z2, H2) = forward(params, X, nin, hidden, nout, bf)
    if bf:
        # the GPU keeps activations in FP32 between the products
        H1, H2, Z1, Z2 = [np.asarray(a, dtype=np.float32).astype(np.float64) for a in (H1, H2, Z1, Z2)]
        Y = np.asarray(Y, dtype=np.float32).astype(np.float64)
    dY = 2.0 * (Y - T) / Y.size
    if bf:
        dY = np.asarray(dY, dtype=np.float32).astype(np.float64)
    gW3 = _r(dY, bf).T @ _r(H2, bf)
    gb3 = dY.sum(0)
    dH2 = _r(dY, bf) @ _r(W3, bf)
    dZ2 = dH2 * dsilu(Z2)
    gW2 = _r(dZ2, bf).T @ _r(H1, bf)
    gb2 = dZ2.sum(0)
    dH1 = _r(dZ2, bf) @ _r(W2, bf)
    dZ1 = dH1 * dsilu(Z1)
    gW1 = _r(dZ1, bf).T @ _r(X, bf)
    gb1 = dZ1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


# --- L2 projection between nodal orders on [-1, 1] --------------------------

def cgl_nodes(p):
    return np.cos((p - np.arange(p + 1)) * np.pi / p)


def lagrange(nodes, x):
    """L[q, i] = l_i(x_q) for the Lagrange basis on `nodes`."""
    x = np.atleast_1d(x)
    L = np.ones((len(x), len(nodes)))
    for i, ri in enumerate(nodes):
        for j, rj in enumerate(nodes):
            if j != i:
                L[:, i] *= (x - rj) / (ri - rj)
    return L


def projection_1d(pf, pc):
    """P (nc x nf): coarse nodal values of the L2 projection of a fine nodal
    polynomial, from a Gauss rule exact for every product involved."""
    g, w = np.polynomial.legendre.leggauss(pf + pc + 2)
    Lf = lagrange(cgl_nodes(pf), g)
    Lc = lagrange(cgl_nodes(pc), g)
    M = Lc.T @ (w[:, None] * Lc)
    B = Lc.T @ (w[:, None] * Lf)
    return np.linalg.solve(M, B)


def project_state(u, pf, pc):
    """u (K, 5, nf^3) -> (K, 5, nc^3), node index ix + n iy + n^2 iz."""
    P = projection_1d(pf, pc)
    nf, nc = pf + 1, pc + 1
    K = u.shape[0]
    a = u.reshape(K, 5, nf, nf, nf)                 # (element, field, iz, iy, ix)
    a = np.einsum("kx,efzyx->efzyk", P, a)          # along x
    a = np.einsum("ky,efzyx->efzkx", P, a)          # along y
    a = np.einsum("kz,efzyx->efkyx", P, a)          # along z
    return a.reshape(K, 5, nc ** 3)
