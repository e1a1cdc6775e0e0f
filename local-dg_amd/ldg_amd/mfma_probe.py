"""Times the rhs kernel's y/z test-function contraction at p = 4 in three
forms (libldg_probe.so, csrc/mfma_probe.hip): the solver's vector sweeps,
FP32-input MFMA and FP16 hi/lo MFMA on the dense 25x25 slab operators, on
64^3 elements of random flux fields, and checks each against an FP64
evaluation of the same contraction.

    python -m ldg_amd.mfma_probe [--elements K] [--reps R]

Prints one JSON line.  Each launch stages 4 elements per workgroup in LDS
and repeats the contraction R times there, in place (every repetition reads
what the one before wrote, as the solver's sweeps overwrite their inputs);
the time of one contraction is (t(R) - t(1)) / (R - 1), which leaves out the
HBM transfer of the fields."""
import argparse
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N = 5


def basis(p=4):
    """1D interpolation I[q][j] = l_j(g_q) and derivative D[q][j] = l_j'(g_q)
    of the Lagrange basis on the Chebyshev-Gauss-Lobatto nodes at the Gauss
    points."""
    n = p + 1
    r = -np.cos(np.pi * np.arange(n) / p)
    g, _ = np.polynomial.legendre.leggauss(n)
    I = np.zeros((n, n))
    D = np.zeros((n, n))
    for j in range(n):
        others = [m for m in range(n) if m != j]
        den = np.prod([r[j] - r[m] for m in others])
        for q in range(n):
            I[q, j] = np.prod([g[q] - r[m] for m in others]) / den
            D[q, j] = sum(np.prod([g[q] - r[m] for m in others if m != k]) for k in others) / den
    return I, D


def symmetrize(A, S):
    return 0.5 * (A + S * A[::-1, ::-1])


def symmat(A):
    """SymMat<5, float>: P[3][2], R[3][2], C[3], A[1][1] (the even/odd split form)."""
    P = np.array([[(A[q, j] + A[q, N - 1 - j]) / 2 for j in range(2)] for q in range(3)])
    R = np.array([[(A[q, j] - A[q, N - 1 - j]) / 2 for j in range(2)] for q in range(3)])
    C = np.array([A[q, 2] for q in range(3)])
    return np.concatenate([P.ravel(), R.ravel(), C, [0.0]]).astype(np.float32)


def operators():
    I, D = basis()
    It = symmetrize(I.T.copy(), 1)
    Dt = symmetrize(D.T.copy(), -1)
    # slab index i = y + 5 z: kron(Z op, Y op)
    Ax = np.kron(It, It)
    Ag = np.hstack([np.kron(It, Dt), np.kron(Dt, It)])
    return It, Dt, Ax, Ag


def pack(It, Dt, Ax, Ag):
    ax = np.zeros((32, 26), np.float32)
    ax[:25, :25] = Ax
    ag = np.zeros((32, 50), np.float32)
    ag[:25, :] = Ag

    def hilo(A, k):
        a = np.zeros((32, k))
        a[:A.shape[0], :A.shape[1]] = A
        h = a.astype(np.float16)
        lo = (a - h.astype(np.float64)).astype(np.float16)
        return h, lo

    axh, axl = hilo(Ax, 32)
    agh, agl = hilo(Ag, 64)
    return b"".join([symmat(It).tobytes(), symmat(Dt).tobytes(), ax.tobytes(), ag.tobytes(),
                     axh.tobytes(), axl.tobytes(), agh.tobytes(), agl.tobytes()])


def reference(F, Ax, Ag, elems):
    """FP64 O_x, O_g of the sampled elements: F (K, 15, 125) -> (len, 10, 125)."""
    out = np.zeros((len(elems), 10, 125))
    for t, e in enumerate(elems):
        f = F[e].astype(np.float64).reshape(15, 25, 5)      # [field][i = y + 5 z][x]
        out[t, :5] = np.einsum("ab,sbx->sax", Ax, f[0:5]).reshape(5, 125)
        out[t, 5:] = (np.einsum("ab,sbx->sax", Ag[:, :25], f[5:10])
                      + np.einsum("ab,sbx->sax", Ag[:, 25:], f[10:15])).reshape(5, 125)
    return out


def measure(elements=64 ** 3, reps=16, samples=64, launches=5, log=None):
    """Times the three variants on `elements` random elements (a multiple of
    4) and checks each against the FP64 contraction on `samples` of them."""
    lib = ctypes.CDLL(os.path.join(HERE, "libldg_probe.so"))
    fp = ctypes.POINTER(ctypes.c_float)
    lib.ldg_probe_measure.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, fp, ctypes.c_long,
                                      ctypes.c_void_p, fp]
    It, Dt, Ax, Ag = operators()
    blob = pack(It, Dt, Ax, Ag)
    if len(blob) != lib.ldg_probe_ops_bytes():
        raise RuntimeError(f"operator block {len(blob)} bytes, the library expects {lib.ldg_probe_ops_bytes()}")
    K = elements
    rng = np.random.default_rng(0)
    F = np.ascontiguousarray(rng.standard_normal((K, 15, 125)).astype(np.float32))
    O = np.empty((K, 10, 125), dtype=np.float32)
    ops = ctypes.create_string_buffer(blob, len(blob))
    elems = rng.choice(K, size=min(samples, K), replace=False)
    ref = reference(F, Ax, Ag, elems)

    def timed(v, r):
        best = ctypes.c_float()
        rc = lib.ldg_probe_measure(v, r, launches, F.ctypes.data_as(fp), O.ctypes.data_as(fp), K, ops,
                                   ctypes.byref(best))
        if rc != 0:
            raise RuntimeError(f"ldg_probe_measure returned {rc}")
        return best.value

    out = {"probe": "y/z test-function contraction of 15 flux fields, p = 4", "elements": K,
           "reps": reps, "variants": {}}
    names = {0: "valu_sum_factorised", 1: "mfma_f32_32x32x2_slab", 2: "mfma_f16_hilo_32x32x16_slab"}
    for v in (0, 1, 2):
        t1 = timed(v, 1)                       # O now holds one contraction
        o = O[elems].astype(np.float64)
        err = float(np.linalg.norm(o - ref) / np.linalg.norm(ref))
        tr = timed(v, reps)
        per = (tr - t1) / (reps - 1)
        out["variants"][names[v]] = {"ms_one_pass_with_hbm": t1, "ms_reps": tr, "ms_per_contraction": per,
                                     "ns_per_element": per * 1e6 / K, "rel_l2_vs_fp64": err}
        if log:
            log(f"{names[v]}: {t1:.4f} ms (1 rep), {tr:.4f} ms ({reps} reps), {per:.4f} ms per contraction, "
                f"error {err:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--elements", type=int, default=64 ** 3)
    ap.add_argument("--reps", type=int, default=16)
    args = ap.parse_args()
    out = measure(args.elements, args.reps, log=lambda m: print(m, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
