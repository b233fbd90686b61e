"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker for the HIP implementation and the
"port" CPU baseline; the product never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
_LIB = None


class OrcParams(C.Structure):
    _fields_ = [("nmin", C.c_int), ("ns", C.c_int), ("nw", C.c_int), ("nibble", C.c_int),
                ("chain", C.c_int), ("maxit", C.c_int)]


class OrcStats(C.Structure):
    _fields_ = [(k, C.c_long) for k in ("iters", "aed_calls", "sweeps", "shifts_used", "ds_iters",
                                        "swaps", "swap_fail", "inf_defl")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(_ROOT, "oracle")])
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.orc_hgeqz.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp,
                                C.c_int, dp, dp, dp, dp, C.c_int, dp, C.c_int, C.POINTER(OrcParams),
                                C.POINTER(OrcStats)]
        L.orc_geqrf_apply.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, C.c_int]
        L.orc_gghrd.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp,
                                C.c_int, C.c_int, C.c_int]
        L.orc_qz_ds.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp, dp, dp, dp,
                                C.c_int, dp, C.c_int, C.c_double, C.POINTER(OrcStats)]
        L.orc_tgex2.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, C.c_int,
                                C.c_int, C.c_int]
        L.orc_tgexc.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_chase_window.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, dp, C.c_int, C.c_int,
                                       C.c_int, C.c_int, dp, dp, dp]
        L.orc_intro_window.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, dp]
        L.orc_tgord.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, C.c_void_p,
                                C.POINTER(C.c_int)]
        L.orc_compres.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, C.c_int, dp]
        L.orc_fill_normal.argtypes = [C.c_uint64, C.c_size_t, dp]
        L.orc_fill_uniform.argtypes = [C.c_uint64, C.c_size_t, dp]
        L.orc_lagv2.argtypes = [dp, C.c_int, dp, C.c_int, dp, dp, dp, dp, dp]
        L.orc_eig22.argtypes = [dp, C.c_int, dp, C.c_int, dp, dp, dp]
        L.orc_shift_col.argtypes = [dp, C.c_int, dp, C.c_int] + [C.c_double] * 6 + [dp]
        L.orc_invhse3.argtypes = [dp, C.c_int, dp, dp]
        L.orc_larfg.argtypes = [C.c_int, dp, dp, C.c_int, dp]
        L.orc_lartg.argtypes = [C.c_double, C.c_double, dp, dp, dp]
        L.orc_mlp_forward.argtypes = [dp, dp, C.c_int, dp]
        L.orc_mlp_grad.argtypes = [dp, dp, dp, C.c_int, dp]
        L.orc_mlp_grad.restype = C.c_double
        for f in ("orc_piparmq_ns", "orc_piparmq_nw", "orc_piparmq_nibble"):
            getattr(L, f).argtypes = [C.c_int]
            getattr(L, f).restype = C.c_int
        _LIB = L
    return _LIB


def ptr(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fmat(a):
    return np.asfortranarray(np.array(a, dtype=np.float64))


def normal(seed, shape):
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_normal(C.c_uint64(seed), n, ptr(out))
    return out.reshape(shape, order="F")


def uniform(seed, shape):
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_uniform(C.c_uint64(seed), n, ptr(out))
    return out.reshape(shape, order="F")


def hgeqz(H, T, job=1, compq=1, compz=1, Q=None, Z=None, ilo=0, ihi=None, params=None):
    """oracle PDHGEQZ: returns (S, P, alphar, alphai, beta, Q, Z, info, stats)."""
    H = fmat(H); T = fmat(T)
    n = H.shape[0]
    ihi = n - 1 if ihi is None else ihi
    Q = fmat(np.eye(n) if Q is None else Q)
    Z = fmat(np.eye(n) if Z is None else Z)
    ar = np.zeros(n); ai = np.zeros(n); be = np.zeros(n)
    prm = params if params is not None else OrcParams()
    st = OrcStats()
    info = lib().orc_hgeqz(job, compq, compz, n, ilo, ihi, ptr(H), n, ptr(T), n, ptr(ar), ptr(ai),
                           ptr(be), ptr(Q), n, ptr(Z), n, C.byref(prm), C.byref(st))
    return H, T, ar, ai, be, Q, Z, info, st.as_dict()


def ht_reduce(A, B, compq=1, compz=1):
    """oracle QR(B) + PDGGHRD: returns (H, T, Q, Z) with H = Q^T A Z, T = Q^T B Z."""
    A = fmat(A); B = fmat(B)
    n = A.shape[0]
    Q = fmat(np.eye(n)); Z = fmat(np.eye(n))
    lib().orc_geqrf_apply(n, ptr(A), n, ptr(B), n, ptr(Q), n, 1 if compq else 0)
    lib().orc_gghrd(n, 0, n - 1, ptr(A), n, ptr(B), n, ptr(Q), n, ptr(Z), n, 2 if compq else 0,
                    2 if compz else 0)
    return A, B, Q, Z


def tgord(S, T, Q, Z, select):
    """oracle reorder (PDTGORD semantics): returns (S, T, Q, Z, select_out, m, info)."""
    S, T, Q, Z = (fmat(x) for x in (S, T, Q, Z))
    n = S.shape[0]
    sel = np.ascontiguousarray(np.asarray(select, dtype=np.int32).copy())
    m = C.c_int(0)
    info = lib().orc_tgord(n, ptr(S), n, ptr(T), n, ptr(Q), n, ptr(Z), n, sel.ctypes.data_as(C.c_void_p),
                           C.byref(m))
    return S, T, Q, Z, sel, m.value, info


def compres(A, B, S, T, Q, Z):
    n = A.shape[0]
    out = np.zeros(4)
    A, B, S, T, Q, Z = (fmat(x) for x in (A, B, S, T, Q, Z))
    lib().orc_compres(n, ptr(A), ptr(B), ptr(S), ptr(T), ptr(Q), ptr(Z), n, ptr(out))
    return out


def mlp_forward(P, X):
    """oracle perceptron forward pass (double precision): X (B x 12) -> Y (B x 3)."""
    P = np.ascontiguousarray(P, dtype=np.float64); X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.zeros((X.shape[0], 3))
    lib().orc_mlp_forward(ptr(P), ptr(X.reshape(-1)), X.shape[0], ptr(Y.reshape(-1)))
    return Y


def mlp_grad(P, X, T):
    """oracle loss and gradient: returns (L, dL/dP)."""
    P = np.ascontiguousarray(P, dtype=np.float64); X = np.ascontiguousarray(X, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    G = np.zeros(P.size)
    Lv = lib().orc_mlp_grad(ptr(P), ptr(X.reshape(-1)), ptr(T.reshape(-1)), X.shape[0], ptr(G))
    return Lv, G
