"""Python surface of the MI355X multishift QZ (ctypes over libgqz.so, include/gqz.h).

Mirrors the PDHGEQZ / PDGGHRD call surface of NLAFET/GEVP-PDHGEQZ for a 1x1 process grid:

    pdhgeqz(job, compq, compz, H, T, Q=None, Z=None, ilo=1, ihi=None, params=None)
        -> (alphar, alphai, beta, info, stats)         (src/pdhgeqz/PDHGEQZ.f:15-31)
    pdgghrd(compq, compz, A, B, Q=None, Z=None, ilo=1, ihi=None) -> info
                                                        (src/ab2ht/PDGGHRD.f:13-14)

Matrices are torch float64 CUDA (HIP) tensors in column-major layout, i.e. the transpose view
of a contiguous tensor (helpers colmajor()/to_colmajor() below), updated in place.  PyTorch is
only the allocator and stream owner here; every operation runs in the HIP kernels of libgqz.so.
The library refuses to run without a GPU: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class GqzParams(C.Structure):
    _fields_ = [("nmin", C.c_int), ("ns", C.c_int), ("nw", C.c_int), ("nibble", C.c_int),
                ("chain", C.c_int), ("maxit", C.c_int), ("use_policy", C.c_int)]


class GqzStats(C.Structure):
    _fields_ = [(k, C.c_long) for k in ("iters", "aed_calls", "sweeps", "shifts_used", "small_calls",
                                        "swaps", "swap_fail", "inf_defl", "chase_windows",
                                        "slab_launches")] + \
               [(k, C.c_double) for k in ("t_total", "t_aed", "t_sweep", "t_small", "t_inf")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GqzTraceRec(C.Structure):
    _fields_ = [("nh", C.c_int), ("ktop", C.c_int), ("kbot", C.c_int), ("ns", C.c_int), ("nw", C.c_int),
                ("nibble", C.c_int), ("nd", C.c_int), ("ndinf", C.c_int), ("swept", C.c_int),
                ("feat", C.c_float * 12), ("t_iter", C.c_double)]


POLICY_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int),
                        C.POINTER(C.c_int), C.POINTER(C.c_int))

EXPORTS = ["gqz_init", "gqz_finalize", "gqz_version", "gqz_stream", "gqz_sync", "gqz_set_policy",
           "gqz_trace_enable", "gqz_trace_get", "gqz_progress", "gqz_overlap", "gqz_overlap_min", "gqz_sched_replay", "gqz_last_error", "gqz_hgeqz_dev", "gqz_hgeqz",
           "gqz_gghrd_dev", "gqz_gghrd", "gqz_geqrb_dev", "gqz_pdhgeqz_host", "gqz_pdgghrd_host",
           "gqz_compres_dev", "gqz_batched_dev", "gqz_mlp_forward_dev", "gqz_mlp_train_step_dev"]


def lib_path():
    return os.path.join(_HERE, "libgqz.so")


def lib():
    """Load libgqz.so (raises when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError("libgqz.so is not built; run __graft_entry__.build() (or "
                           "python gevp-pdhgeqz_amd/build.py) first")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.gqz_init.argtypes = [C.c_int]
    L.gqz_version.restype = C.c_char_p
    L.gqz_last_error.restype = C.c_char_p
    L.gqz_stream.restype = C.c_void_p
    L.gqz_hgeqz_dev.argtypes = [C.c_char, C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp,
                                C.c_int, vp, vp, vp, vp, C.c_int, vp, C.c_int, C.POINTER(GqzParams),
                                C.POINTER(GqzStats)]
    L.gqz_hgeqz.argtypes = L.gqz_hgeqz_dev.argtypes
    L.gqz_gghrd_dev.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int,
                                vp, C.c_int, vp, C.c_int]
    L.gqz_gghrd.argtypes = L.gqz_gghrd_dev.argtypes
    L.gqz_geqrb_dev.argtypes = [C.c_char, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int]
    L.gqz_compres_dev.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, C.c_int, dp]
    L.gqz_set_policy.argtypes = [POLICY_FN, vp]
    L.gqz_trace_enable.argtypes = [C.c_int]
    L.gqz_trace_get.argtypes = [C.POINTER(GqzTraceRec), C.c_int]
    L.gqz_progress.argtypes = [C.c_double]
    L.gqz_overlap.argtypes = [C.c_int]
    L.gqz_overlap_min.argtypes = [C.c_int]
    L.gqz_sched_replay.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int]
    L.gqz_slab_update_dev.argtypes = [C.c_char, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int]
    L.gqz_chase_window_dev.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.gqz_tgex2_dev.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp]
    L.gqz_small_qz_dev.argtypes = [C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_double]
    L.gqz_batched_dev.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int]
    L.gqz_tgord_dev.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp,
                                C.c_int, vp, vp, vp, ip]
    L.gqz_pdtgord_host.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                   ip, vp, C.c_int, vp, C.c_int]
    L.gqz_mm_info.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, ip, ip, C.POINTER(C.c_long)]
    L.gqz_mm_read_dense.argtypes = [C.c_char_p, C.c_int, C.c_int, vp, C.c_int]
    L.gqz_batched_noalloc_dev.argtypes = L.gqz_batched_dev.argtypes
    L.gqz_batched_tune.argtypes = [C.c_int, vp, ip, ip, dp]
    _LIB = L
    return L


def init(device=0):
    L = lib()
    if L.gqz_init(device) != 0:
        raise RuntimeError("gqz_init failed: " + L.gqz_last_error().decode())
    return L


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def colmajor(n, m=None, device="cuda"):
    """An n x m float64 column-major device matrix (a transposed view of a contiguous tensor)."""
    import torch
    m = n if m is None else m
    return torch.zeros((m, n), dtype=torch.float64, device=device).t()


def to_colmajor(a, device="cuda"):
    import torch
    t = torch.as_tensor(a, dtype=torch.float64)
    return t.t().contiguous().to(device).t()


def ld(t):
    """Leading dimension of a column-major matrix view."""
    assert t.stride(0) == 1, "matrix must be column-major (use colmajor()/to_colmajor())"
    return t.stride(1) if t.dim() == 2 and t.shape[1] > 1 else max(1, t.shape[0])


def _sync_torch():
    import torch
    torch.cuda.synchronize()


def pdhgeqz(job, compq, compz, H, T, Q=None, Z=None, ilo=1, ihi=None, params=None, alphar=None,
            alphai=None, beta=None):
    """Multishift QZ on device matrices (PDHGEQZ semantics, 1x1 grid).  Returns
    (alphar, alphai, beta, info, stats) with eigenvalue arrays on the device."""
    import torch
    L = init()
    n = H.shape[0]
    ihi = n if ihi is None else ihi
    dev = H.device
    ar = torch.empty(n, dtype=torch.float64, device=dev) if alphar is None else alphar
    ai = torch.empty(n, dtype=torch.float64, device=dev) if alphai is None else alphai
    be = torch.empty(n, dtype=torch.float64, device=dev) if beta is None else beta
    prm = params if params is not None else GqzParams()
    st = GqzStats()
    _sync_torch()
    info = L.gqz_hgeqz_dev(job.encode()[:1], compq.encode()[:1], compz.encode()[:1], n, ilo, ihi,
                           _p(H), ld(H), _p(T), ld(T), _p(ar), _p(ai), _p(be),
                           _p(Q), ld(Q) if Q is not None else 1, _p(Z), ld(Z) if Z is not None else 1,
                           C.byref(prm), C.byref(st))
    if info == -100:
        raise RuntimeError("gqz_hgeqz_dev failed: " + L.gqz_last_error().decode())
    return ar, ai, be, info, st.as_dict()


def pdgghrd(compq, compz, A, B, Q=None, Z=None, ilo=1, ihi=None):
    """HT reduction of (A, B), B upper triangular (PDGGHRD semantics, PDHGEQZ's Q convention:
    A_out = Q^T A_in Z)."""
    L = init()
    n = A.shape[0]
    ihi = n if ihi is None else ihi
    _sync_torch()
    info = L.gqz_gghrd_dev(compq.encode()[:1], compz.encode()[:1], n, ilo, ihi, _p(A), ld(A), _p(B), ld(B),
                           _p(Q), ld(Q) if Q is not None else 1, _p(Z), ld(Z) if Z is not None else 1)
    if info == -100:
        raise RuntimeError("gqz_gghrd_dev failed: " + L.gqz_last_error().decode())
    return info


def geqrb(compq, A, B, Q=None):
    """QR of B on the device applied to A (and accumulated into Q): B <- R, A <- Q_B^T A."""
    L = init()
    n = A.shape[0]
    _sync_torch()
    info = L.gqz_geqrb_dev(compq.encode()[:1], n, _p(A), ld(A), _p(B), ld(B), _p(Q),
                           ld(Q) if Q is not None else 1)
    if info == -100:
        raise RuntimeError("gqz_geqrb_dev failed: " + L.gqz_last_error().decode())
    return info


def slab_update(side, Cm, F):
    """Kernel-level hook: side 'R': C <- C F ; 'L': C <- F^T C (in place, FP64 MFMA)."""
    L = init()
    k = F.shape[0]
    m = Cm.shape[0] if side == "R" else Cm.shape[1]
    _sync_torch()
    r = L.gqz_slab_update_dev(side.encode()[:1], m, k, _p(Cm), ld(Cm), _p(F), ld(F))
    if r != 0:
        raise RuntimeError("gqz_slab_update_dev failed (%d): %s" % (r, L.gqz_last_error().decode()))


def chase_window(H, T, p, w, nb, t0, steps, intro, last, U, V, sr, si, sb):
    """Kernel-level hook: one window of the chain chase (see include/gqz.h)."""
    L = init()
    _sync_torch()
    r = L.gqz_chase_window_dev(_p(H), _p(T), ld(H), p, w, nb, t0, steps, intro, last, _p(U), _p(V),
                               _p(sr), _p(si), _p(sb))
    if r != 0:
        raise RuntimeError("gqz_chase_window_dev failed (%d): %s" % (r, L.gqz_last_error().decode()))


def tgex2(n1, n2, S, T):
    """Kernel-level hook: one DTGEX2 swap of the (n1+n2)-order corner (S, T) (column-major device
    matrices).  Returns (rejected, Qa, Za, S', T') as host numpy arrays."""
    import torch
    L = init()
    m = n1 + n2
    out = torch.zeros(65, dtype=torch.float64, device=S.device)
    _sync_torch()
    r = L.gqz_tgex2_dev(n1, n2, _p(S), _p(T), ld(S), _p(out))
    if r != 0:
        raise RuntimeError("gqz_tgex2_dev failed (%d): %s" % (r, L.gqz_last_error().decode()))
    o = out.cpu().numpy()
    blk = lambda k: o[1 + 16 * k:1 + 16 * k + m * m].reshape(m, m, order="F")
    return bool(o[0] != 0.0), blk(0), blk(1), blk(2), blk(3)


def small_qz(H, T, U, V, alphar, alphai, beta, ttol):
    """Kernel-level hook: LDS-resident base QZ of the whole (small) pencil; returns INFO."""
    L = init()
    m = H.shape[0]
    _sync_torch()
    return L.gqz_small_qz_dev(m, _p(H), _p(T), ld(H), _p(U), _p(V), _p(alphar), _p(alphai), _p(beta),
                              float(ttol))


def tgord(select, S, T, Q=None, Z=None, para=None):
    """Reorder the generalized Schur form (S, T) (device, column-major) so that the eigenvalues with
    select[j] != 0 lead (PDTGORD, src/reorder/PDTGORD.f:1-4); Q, Z (device) are postmultiplied when
    given.  Returns (info, m, select_out, alphar, alphai, beta) with the eigenvalues on the device."""
    import numpy as np
    import torch
    L = init()
    n = S.shape[0]
    sel = np.ascontiguousarray(np.asarray(select, dtype=np.int32).copy())
    pa = None if para is None else np.ascontiguousarray(para, dtype=np.int32)
    ar = torch.empty(n, dtype=torch.float64, device=S.device); ai = torch.empty_like(ar); be = torch.empty_like(ar)
    m = C.c_int(0)
    _sync_torch()
    info = L.gqz_tgord_dev(int(Q is not None), int(Z is not None), sel.ctypes.data_as(C.c_void_p),
                           None if pa is None else pa.ctypes.data_as(C.c_void_p), n, _p(S), ld(S), _p(T), ld(T),
                           _p(Q), ld(Q) if Q is not None else 1, _p(Z), ld(Z) if Z is not None else 1,
                           _p(ar), _p(ai), _p(be), C.byref(m))
    if info == -100:
        raise RuntimeError("gqz_tgord_dev failed: " + L.gqz_last_error().decode())
    return info, m.value, sel, ar, ai, be


def mm_read(path):
    """Read a Matrix Market file into a dense column-major numpy array with the library's reader
    (mminfo2 / PDRDMM roles); a .gz file is decompressed to a temporary file first.  Returns
    (A, info_dict)."""
    import gzip
    import shutil
    import tempfile
    import numpy as np
    L = lib()
    tmp = None
    try:
        if path.endswith(".gz"):
            fd, tmp = tempfile.mkstemp(suffix=".mtx")
            with os.fdopen(fd, "wb") as out, gzip.open(path, "rb") as src:
                shutil.copyfileobj(src, out)
            path = tmp
        rep, field, symm = C.create_string_buffer(16), C.create_string_buffer(16), C.create_string_buffer(24)
        m, n, nnz = C.c_int(0), C.c_int(0), C.c_long(0)
        r = L.gqz_mm_info(path.encode(), rep, field, symm, C.byref(m), C.byref(n), C.byref(nnz))
        if r != 0:
            raise RuntimeError("gqz_mm_info(%s) returned %d" % (path, r))
        A = np.zeros((m.value, n.value), order="F")
        r = L.gqz_mm_read_dense(path.encode(), m.value, n.value, A.ctypes.data_as(C.c_void_p), max(1, m.value))
        if r != 0:
            raise RuntimeError("gqz_mm_read_dense(%s) returned %d" % (path, r))
        return A, {"rep": rep.value.decode(), "field": field.value.decode(), "symm": symm.value.decode(),
                   "nnz": nnz.value}
    finally:
        if tmp:
            os.unlink(tmp)


def compres(A, B, S, T, Q, Z):
    """PCOMPRES residuals on the device: [resA, resB, resQ/(n eps), resZ/(n eps)]."""
    L = init()
    n = A.shape[0]
    out = (C.c_double * 4)()
    _sync_torch()
    r = L.gqz_compres_dev(n, _p(A), _p(B), _p(S), _p(T), _p(Q), _p(Z), ld(A), out)
    if r != 0:
        raise RuntimeError("gqz_compres_dev failed: " + L.gqz_last_error().decode())
    return list(out)


def schur_pipeline(A, B, params=None):
    """Full pipeline on device copies: QR(B), HT reduction, QZ with Q and Z accumulated.
    Returns (S, P, Q, Z, alphar, alphai, beta, info, stats)."""
    n = A.shape[0]
    S = A.clone() if A.stride(0) == 1 else to_colmajor(A.cpu())
    P = B.clone() if B.stride(0) == 1 else to_colmajor(B.cpu())
    S = S.t().contiguous().t(); P = P.t().contiguous().t()
    Q = colmajor(n, device=A.device); Z = colmajor(n, device=A.device)
    geqrb("I", S, P, Q)
    pdgghrd("V", "I", S, P, Q, Z)
    ar, ai, be, info, st = pdhgeqz("S", "V", "V", S, P, Q, Z, params=params)
    return S, P, Q, Z, ar, ai, be, info, st
