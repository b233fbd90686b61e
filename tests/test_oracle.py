"""CPU tests: pin the oracle (oracle/qz_oracle.c) against LAPACK, closed forms, constructed
pencils with known spectra and the committed golden fixtures.  No GPU needed."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle_ffi as O
import qzcheck as Q

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


def test_generator_fixture():
    """SplitMix64 + Box-Muller generator: first values and checksums match the fixture."""
    g = np.load(GOLD)
    np.testing.assert_array_equal(O.normal(1, (16,)), g["normal_seed1_16"])
    np.testing.assert_array_equal(O.uniform(7, (16,)), g["uniform_seed7_16"])
    x = O.normal(12345, (100000,))
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1.0) < 0.01


def test_piparmq_table():
    """PIPARMQ defaults (piparmq.f:445-517) at the config sizes: NS/NW/NIBBLE = 256/384/7 at
    n=8000, 512/768/6 at 12000, 512/768/5 at 16000 (UserGuide p.13)."""
    L = O.lib()
    assert (L.orc_piparmq_ns(8000), L.orc_piparmq_nw(8000), L.orc_piparmq_nibble(8000)) == (256, 384, 7)
    assert (L.orc_piparmq_ns(12000), L.orc_piparmq_nw(12000), L.orc_piparmq_nibble(12000)) == (512, 768, 6)
    assert (L.orc_piparmq_ns(16000), L.orc_piparmq_nw(16000), L.orc_piparmq_nibble(16000)) == (512, 768, 5)
    assert L.orc_piparmq_ns(100) == 10 and L.orc_piparmq_nw(100) == 10
    assert L.orc_piparmq_nw(1000) == 96 and L.orc_piparmq_nibble(1000) == 14


def test_lartg_larfg():
    import ctypes as C
    L = O.lib()
    c, s, r = C.c_double(), C.c_double(), C.c_double()
    for f, g in [(3.0, 4.0), (-1.0, 2.0), (0.0, -5.0), (7.0, 0.0)]:
        L.orc_lartg(f, g, C.byref(c), C.byref(s), C.byref(r))
        assert abs(c.value * f + s.value * g - r.value) < 1e-15 * abs(r.value) + 1e-300
        assert abs(-s.value * f + c.value * g) < 1e-15 * max(abs(f), abs(g))
        assert abs(c.value ** 2 + s.value ** 2 - 1.0) < 1e-15
    x = np.array([2.0, -1.0, 0.5])
    alpha = C.c_double(x[0]); tail = x[1:].copy(); tau = C.c_double()
    L.orc_larfg(3, C.byref(alpha), O.ptr(tail), 1, C.byref(tau))
    v = np.r_[1.0, tail]
    Hm = np.eye(3) - tau.value * np.outer(v, v)
    np.testing.assert_allclose(Hm @ x, [alpha.value, 0, 0], atol=1e-15)
    np.testing.assert_allclose(Hm @ Hm.T, np.eye(3), atol=1e-15)


def test_shift_column_closed_form():
    """QZFCOL role: v is parallel to p(H T^-1) e1 with p(z) = (z - l1)(z - l2) (closed form)."""
    rng = np.random.default_rng(3)
    for cplx in (False, True):
        H = np.triu(rng.standard_normal((4, 4)), -1)
        T = np.triu(rng.standard_normal((4, 4))) + 3 * np.eye(4)
        if cplx:
            l1, l2 = 0.3 + 0.7j, 0.3 - 0.7j
            args = (0.3, 0.7, 1.0, 0.3, -0.7, 1.0)
        else:
            l1, l2 = 0.4, -1.1
            args = (0.4, 0.0, 1.0, -1.1, 0.0, 1.0)
        M = H @ np.linalg.inv(T)
        p = (M - l1 * np.eye(4)) @ (M - l2 * np.eye(4)) @ np.eye(4)[:, 0]
        p = np.real(p)
        v = np.zeros(3)
        O.lib().orc_shift_col(O.ptr(O.fmat(H)), 4, O.ptr(O.fmat(T)), 4, *args, O.ptr(v))
        assert abs(p[3]) < 1e-12
        cosang = abs(v @ p[:3]) / (np.linalg.norm(v) * np.linalg.norm(p[:3]))
        assert abs(cosang - 1.0) < 1e-13


def test_invhse_annihilates():
    """INVHSE: T (I - tau v v^T) has zeros below the diagonal in its first column."""
    rng = np.random.default_rng(5)
    for _ in range(5):
        T = rng.standard_normal((3, 3))
        v = np.zeros(3); tau = np.zeros(1)
        O.lib().orc_invhse3(O.ptr(O.fmat(T)), 3, O.ptr(v), O.ptr(tau))
        P = np.eye(3) - tau[0] * np.outer(v, v)
        R = T @ P
        assert np.abs(R[1:, 0]).max() < 1e-14 * np.abs(T).max()
        np.testing.assert_allclose(P @ P.T, np.eye(3), atol=1e-15)


def test_lagv2_2x2_forms():
    """DLAGV2 role: real pair -> triangular; complex pair -> B diagonal positive; same eigenvalues."""
    rng = np.random.default_rng(9)
    for trial in range(40):
        A = rng.standard_normal((2, 2))
        B = np.triu(rng.standard_normal((2, 2))) + np.diag([0.5, 0.5]) * np.sign(rng.standard_normal(2))
        A0, B0 = A.copy(), B.copy()
        Af, Bf = O.fmat(A), O.fmat(B)
        ar, ai, be = np.zeros(2), np.zeros(2), np.zeros(2)
        QL, ZR = np.zeros(4), np.zeros(4)
        O.lib().orc_lagv2(O.ptr(Af), 2, O.ptr(Bf), 2, O.ptr(ar), O.ptr(ai), O.ptr(be), O.ptr(QL), O.ptr(ZR))
        QLm, ZRm = QL.reshape(2, 2, order="F"), ZR.reshape(2, 2, order="F")
        np.testing.assert_allclose(QLm.T @ A0 @ ZRm, Af, atol=1e-13 * (1 + np.abs(A0).max()))
        np.testing.assert_allclose(QLm.T @ B0 @ ZRm, Bf, atol=1e-13 * (1 + np.abs(B0).max()))
        ref = sl.eigvals(A0, B0)
        got = (ar + 1j * ai) / be
        d = min(np.abs(got - ref).max(), np.abs(got[::-1] - ref).max())
        assert d < 1e-10 * (1 + np.abs(ref).max())
        if ai[0] == 0:
            assert Af[1, 0] == 0.0 and Bf[1, 0] == 0.0
        else:
            assert Bf[0, 1] == 0.0 and Bf[0, 0] > 0 and Bf[1, 1] > 0


@pytest.mark.parametrize("n1,n2", [(1, 1), (1, 2), (2, 1), (2, 2)])
def test_tgex2_swaps_blocks(n1, n2):
    """DTGEX2 role: swapping two blocks keeps the pencil equivalent and exchanges eigenvalues."""
    rng = np.random.default_rng(10 * n1 + n2)
    m = n1 + n2

    def block(k):
        if k == 1:
            return np.array([[rng.uniform(1, 2)]]), np.array([[rng.uniform(0.5, 1)]])
        a, b = rng.uniform(0.5, 1.5), rng.uniform(1, 2)
        return np.array([[a, b], [-b, a]]), np.diag(rng.uniform(0.5, 1.5, 2))

    S11, T11 = block(n1); S22, T22 = block(n2)
    S = np.zeros((m, m)); T = np.zeros((m, m))
    S[:n1, :n1], T[:n1, :n1] = S11, T11
    S[n1:, n1:], T[n1:, n1:] = S22, T22
    S[:n1, n1:] = rng.standard_normal((n1, n2)); T[:n1, n1:] = rng.standard_normal((n1, n2))
    S0, T0 = S.copy(), T.copy()
    Sf, Tf, Qf, Zf = O.fmat(S), O.fmat(T), O.fmat(np.eye(m)), O.fmat(np.eye(m))
    r = O.lib().orc_tgex2(m, O.ptr(Sf), m, O.ptr(Tf), m, O.ptr(Qf), m, O.ptr(Zf), m, 0, n1, n2)
    assert r == 0
    np.testing.assert_allclose(Qf.T @ S0 @ Zf, Sf, atol=1e-13)
    np.testing.assert_allclose(Qf.T @ T0 @ Zf, Tf, atol=1e-13)
    assert np.abs(Sf[n2:, :n2]).max() == 0.0 and np.abs(np.tril(Tf, -1)).max() == 0.0
    top_ref = sl.eigvals(S22, T22)
    top_got = sl.eigvals(Sf[:n2, :n2], Tf[:n2, :n2])
    d = min(np.abs(top_got - top_ref).max(), np.abs(top_got[::-1] - top_ref).max())
    assert d < 1e-12 * (1 + np.abs(top_ref).max())


def test_ht_reduction_invariants():
    """QR(B) + PDGGHRD: H upper Hessenberg, T upper triangular, backward error and
    orthogonality at PCOMPRES thresholds (EXRAND2.f:404-407)."""
    A, B = Q.pencil_randn(60, 1)
    H, T, Qm, Zm = O.ht_reduce(A, B)
    assert np.abs(np.tril(H, -2)).max() == 0.0 and np.abs(np.tril(T, -1)).max() == 0.0
    ra, rb, rq, rz = Q.compres(A, B, H, T, Qm, Zm)
    assert ra < 1e-14 and rb < 1e-14 and rq < 10 and rz < 10


@pytest.mark.parametrize("n", [1, 2, 3, 7, 40, 120, 260])
def test_oracle_qz_vs_lapack(n):
    """Oracle multishift QZ with AED: PCOMPRES thresholds, Schur structure, and the eigenvalue set
    equal to LAPACK DGGES's within the condition-weighted chordal bound."""
    A, B = Q.pencil_randn(n, 100 + n)
    H, T, Qm, Zm = O.ht_reduce(A, B)
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=2, compz=2, Q=Qm, Z=Zm)
    assert info == 0
    ra, rb, rq, rz = Q.compres(A, B, S, P, Q2, Z2)
    assert ra < 1e-14 and rb < 1e-14 and rq < 10 and rz < 10
    below, tb, ok2 = Q.schur_structure(S, P)
    assert below == 0.0 and tb == 0.0 and ok2
    assert (be >= 0).all()
    ratio, _, _ = Q.match_eigs(ar, ai, be, A, B)
    assert ratio < 1.0


def test_oracle_n1_n2_closed_forms():
    """n = 1: (alpha, beta) = (a, |b|) with sign; n = 2 eigenvalues from the quadratic."""
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(np.array([[3.0]]), np.array([[-2.0]]))
    assert info == 0 and be[0] == 2.0 and ar[0] == -3.0
    A = np.array([[1.0, 2.0], [3.0, 4.0]]); B = np.array([[2.0, 1.0], [0.0, 1.0]])
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(A, B)
    # det(A - l B) = (1 - 2l)(4 - l) - (2 - l) 3 = 2 l^2 - 6 l - 2
    roots = np.sort(np.roots([2.0, -6.0, -2.0]))
    np.testing.assert_allclose(np.sort(ar / be), roots, rtol=1e-13)


@pytest.mark.parametrize("n,k", [(40, 4), (150, 12), (90, 45)])
def test_oracle_known_infinite_count(n, k):
    """Constructed pencil with exactly k semisimple infinite eigenvalues: beta = 0 count == k and
    the finite eigenvalues equal the constructed ones."""
    A, B, lam, kinf = Q.pencil_known(n, 7 + n, n_inf=k)
    H, T, Qm, Zm = O.ht_reduce(A, B)
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=2, compz=2, Q=Qm, Z=Zm)
    assert info == 0
    assert int((be == 0).sum()) == kinf
    fin = be != 0
    got = np.sort((ar[fin] + 1j * ai[fin]) / be[fin])
    np.testing.assert_allclose(np.sort_complex(got), np.sort_complex(lam.astype(complex)), rtol=1e-8, atol=1e-8)


def test_oracle_vs_golden_lapack():
    """The committed golden fixture (LAPACK DGGES eigenvalues + residual bounds for seeded
    pencils, written by tests/golden/make_golden.py) is reproduced by the oracle."""
    g = np.load(GOLD)
    for n in (16, 48):
        A, B = Q.pencil_randn(n, int(g["seed_%d" % n]))
        np.testing.assert_array_equal(A[:2, :2], g["A_%d_corner" % n])
        H, T, Qm, Zm = O.ht_reduce(A, B)
        S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=2, compz=2, Q=Qm, Z=Zm)
        a2, b2 = g["lapack_alpha_%d" % n], g["lapack_beta_%d" % n]
        D = Q.chordal((ar + 1j * ai)[:, None], be[:, None], a2[None, :], b2[None, :])
        from scipy.optimize import linear_sum_assignment
        ri, ci = linear_sum_assignment(D)
        assert D[ri, ci].max() < 1e-10


def test_oracle_chase_window_equivalence():
    """PDHGEQZB chase of a chain equals successive applications (invariants: orthogonal U, V,
    U^T H0 V = H1, Hessenberg-with-bulges structure after push-off)."""
    n = 40
    H0, T0 = Q.pencil_ht_random(n, 2)
    # build a chain of 3 bulges with the oracle's own intro via one double-shift sweep each
    H, T = O.fmat(H0), O.fmat(T0)
    U = O.fmat(np.eye(n)); V = O.fmat(np.eye(n))
    sr = np.array([0.5, -0.3, 1.2, 0.7, -1.0, 0.2]); si = np.zeros(6); sb = np.ones(6)
    # chain already at tails 0,3,6 is not needed for push-off invariants: use movedist = n
    O.lib().orc_chase_window(n, O.ptr(H), n, O.ptr(T), n, O.ptr(U), O.ptr(V), n, 1, n, 1,
                             O.ptr(sr), O.ptr(si), O.ptr(sb))
    np.testing.assert_allclose(U.T @ U, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(U.T @ H0 @ V, H, atol=1e-12 * np.abs(H0).max())
    np.testing.assert_allclose(U.T @ T0 @ V, T, atol=1e-12 * np.abs(T0).max())


# ---- perceptron of the learned parameter policy ---------------------------------------------------
def _mlp_case(seed, B=7):
    rng = np.random.default_rng(seed)
    P = rng.standard_normal(5187) * 0.2
    X = rng.standard_normal((B, 12))
    T = rng.standard_normal((B, 3))
    return P, X, T


def test_mlp_oracle_gradient_matches_finite_differences():
    """The oracle's analytic gradient (pinned here) against central differences of its loss, on a
    sample of parameters from every block (W1, b1, W2, b2, W3, b3)."""
    P, X, T = _mlp_case(3)
    L0, G = O.mlp_grad(P, X, T)
    loss = lambda Q: float(np.mean((O.mlp_forward(Q, X) - T) ** 2))
    assert abs(L0 - loss(P)) < 1e-12 * max(1.0, L0)
    rng = np.random.default_rng(4)
    blocks = [(0, 768), (768, 832), (832, 4928), (4928, 4992), (4992, 5184), (5184, 5187)]
    idx = np.concatenate([rng.integers(a, b, size=min(6, b - a)) for a, b in blocks])
    h = 1e-6
    for i in idx:
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        fd = (loss(Pp) - loss(Pm)) / (2 * h)
        assert abs(fd - G[i]) <= 1e-6 * max(1.0, abs(G[i])), (i, fd, G[i])


def test_mlp_oracle_forward_closed_form():
    """With W2 = 0 the network output is W3 gelu(b2) + b3 for every input (a closed form)."""
    P, X, _ = _mlp_case(5)
    P[832:4928] = 0.0
    from math import erf, sqrt
    gelu = lambda z: 0.5 * z * (1 + erf(z / sqrt(2)))
    a2 = np.array([gelu(z) for z in P[4928:4992]])
    W3 = P[4992:5184].reshape(3, 64)
    expect = W3 @ a2 + P[5184:5187]
    Y = O.mlp_forward(P, X)
    assert np.allclose(Y, expect[None, :], rtol=0, atol=1e-13)


def _max_matched_rel_err(got, ref):
    """Largest |got - ref| / (1 + |ref|) over the best one-to-one matching of two eigenvalue lists
    (sorting alone can split a conjugate pair whose real parts differ in the last bit)."""
    from scipy.optimize import linear_sum_assignment
    D = np.abs(got[:, None] - ref[None, :]) / (1.0 + np.abs(ref[None, :]))
    ri, ci = linear_sum_assignment(D)
    return D[ri, ci].max()


def _mp_gen_eigs(A, B, dps=50):
    """Roots of det(A - l B) by brute force in 50-digit arithmetic: the degree-n polynomial is
    interpolated from n + 1 exact-enough determinant values and solved with mpmath.polyroots."""
    import mpmath as mp
    mp.mp.dps = dps
    n = A.shape[0]
    Am, Bm = mp.matrix(A.tolist()), mp.matrix(B.tolist())
    xs = [mp.mpf(k) for k in range(n + 1)]
    ys = [mp.det(Am - x * Bm) for x in xs]
    # Newton divided differences -> monomial coefficients (highest degree first)
    c = list(ys)
    for j in range(1, n + 1):
        for i in range(n, j - 1, -1):
            c[i] = (c[i] - c[i - 1]) / (xs[i] - xs[i - j])
    poly = [mp.mpf(0)] * (n + 1)      # ascending powers
    poly[0] = c[n]
    for k in range(n - 1, -1, -1):    # poly <- poly * (x - xs[k]) + c[k]
        new = [mp.mpf(0)] * (n + 1)
        for i in range(n):
            new[i + 1] += poly[i]
            new[i] -= xs[k] * poly[i]
        new[0] += c[k]
        poly = new
    roots = mp.polyroots(list(reversed(poly)), maxsteps=200, extraprec=200)
    return np.array([complex(r) for r in roots])


@pytest.mark.parametrize("n,seed", [(3, 1), (4, 2), (5, 3), (6, 4), (6, 5)])
def test_oracle_tiny_pencils_brute_force(n, seed):
    """n <= 6: the oracle's eigenvalues equal the roots of det(H - l T) found by brute force in
    50-digit arithmetic (no LAPACK involved)."""
    H, T = Q.pencil_ht_random(n, seed)
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T)
    assert info == 0 and (be > 0).all()
    assert _max_matched_rel_err((ar + 1j * ai) / be, _mp_gen_eigs(H, T)) < 1e-9


def test_oracle_triangular_pencil_is_exact():
    """A pencil that is already upper triangular deflates at once: alpha_j = +-H(j,j) and
    beta_j = |T(j,j)| exactly, in diagonal order, with beta = 0 exactly where T(j,j) = 0."""
    n = 30
    rng = np.random.default_rng(11)
    H = np.triu(rng.standard_normal((n, n)))
    T = np.triu(rng.standard_normal((n, n)))
    T[[4, 17], [4, 17]] = 0.0
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=1, compz=1)
    assert info == 0
    assert np.all(ai == 0.0)
    np.testing.assert_array_equal(be, np.abs(np.diag(T)))
    fin = np.diag(T) != 0
    np.testing.assert_array_equal(ar[fin], np.sign(np.diag(T))[fin] * np.diag(H)[fin])
    assert int((be == 0).sum()) == 2


def test_oracle_standard_problem_b_identity():
    """B = I reduces the pencil to the standard eigenproblem: the oracle's eigenvalues equal
    LAPACK DHSEQR's (numpy eigvals) of the Hessenberg H within a condition-weighted bound."""
    n = 80
    H, _ = Q.pencil_ht_random(n, 21)
    T = np.eye(n)
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=1, compz=1)
    assert info == 0
    ra, rb, rq, rz = Q.compres(H, T, S, P, Q2, Z2)
    assert ra < 1e-14 and rb < 1e-14 and rq < 10 and rz < 10
    ratio, _, _ = Q.match_eigs(ar, ai, be, H, T)
    assert ratio < 1.0
    assert _max_matched_rel_err((ar + 1j * ai) / be, np.linalg.eigvals(H)) < 1e-8


def test_compres_two_implementations_agree():
    """The oracle's PCOMPRES (C) and the numpy restatement agree on a QZ result, and both see a
    transposed Q as a backward-error failure."""
    A, B = Q.pencil_randn(50, 31)
    H, T, Qm, Zm = O.ht_reduce(A, B)
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=2, compz=2, Q=Qm, Z=Zm)
    c = O.compres(A, B, S, P, Q2, Z2)
    p = np.array(Q.compres(A, B, S, P, Q2, Z2))
    assert c[0] < 1e-14 and c[1] < 1e-14
    # the residuals are themselves rounding errors: two summation orders agree to a few percent
    np.testing.assert_allclose(c[2:], p[2:], rtol=0.05)
    np.testing.assert_allclose(c[:2], p[:2], rtol=0.2, atol=2e-16)
    bad = O.compres(A, B, S, P, Q2.T.copy(), Z2)
    assert bad[0] > 1e-3 and Q.compres(A, B, S, P, Q2.T.copy(), Z2)[0] > 1e-3


def test_base_double_shift_qz():
    """orc_qz_ds (the small-block QZ) on its own: Schur form of an HT pencil with the PCOMPRES
    bounds and LAPACK's eigenvalues."""
    n = 30
    H0, T0 = Q.pencil_ht_random(n, 5)
    H, T = O.fmat(H0.copy()), O.fmat(T0.copy())
    Qm, Zm = O.fmat(np.eye(n)), O.fmat(np.eye(n))
    ar, ai, be = np.zeros(n), np.zeros(n), np.zeros(n)
    ttol = np.finfo(float).eps * np.linalg.norm(T0)
    info = O.lib().orc_qz_ds(n, 0, n - 1, O.ptr(H), n, O.ptr(T), n, O.ptr(ar), O.ptr(ai), O.ptr(be),
                             O.ptr(Qm), n, O.ptr(Zm), n, ttol, None)
    assert info == 0
    ra, rb, rq, rz = Q.compres(H0, T0, H, T, Qm, Zm)
    assert ra < 1e-14 and rb < 1e-14 and rq < 10 and rz < 10
    below, tb, ok2 = Q.schur_structure(H, T)
    assert below == 0.0 and tb == 0.0 and ok2
    ratio, _, _ = Q.match_eigs(ar, ai, be, H0, T0)
    assert ratio < 1.0


def test_tgexc_moves_a_block_to_the_top():
    """DTGEXC semantics (src/reorder/dtgexc.f): a real pencil with real eigenvalues in Schur form;
    moving the last 1x1 block to the top puts its eigenvalue first and shifts the others down by one,
    with Q, Z updated so that the backward error stays at rounding level."""
    n = 12
    A, B, lam, _ = Q.pencil_known(n, 17)
    H, T, Qm, Zm = O.ht_reduce(A, B)
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=2, compz=2, Q=Qm, Z=Zm)
    assert info == 0 and np.all(ai == 0.0)
    before = np.diag(S) / np.diag(P)
    ifst, ilst = C.c_int(n - 1), C.c_int(0)
    S, P, Q2, Z2 = (O.fmat(x) for x in (S, P, Q2, Z2))
    r = O.lib().orc_tgexc(n, O.ptr(S), n, O.ptr(P), n, O.ptr(Q2), n, O.ptr(Z2), n, C.byref(ifst), C.byref(ilst))
    assert r == 0
    after = np.diag(S) / np.diag(P)
    np.testing.assert_allclose(after, np.concatenate([[before[-1]], before[:-1]]), rtol=1e-10)
    assert np.abs(np.tril(S, -1)).max() < 1e-13 * np.abs(S).max() and np.abs(np.tril(P, -1)).max() == 0.0
    ra, rb, rq, rz = Q.compres(A, B, S, P, Q2, Z2)
    assert ra < 1e-14 and rb < 1e-14 and rq < 10 and rz < 10


@pytest.mark.parametrize("A,B", [
    ([[1.0, 2.0], [3.0, 4.0]], [[2.0, 1.0], [0.0, 1.0]]),        # real pair
    ([[1.0, -5.0], [2.0, 1.0]], [[1.0, 0.5], [0.0, 2.0]]),       # complex pair
    ([[0.0, 1.0], [-1.0, 0.0]], [[1.0, 0.0], [0.0, 1.0]]),       # +-i
])
def test_eig22_quadratic_roots(A, B):
    """Eigenvalues of a 2x2 pencil with B upper triangular are the roots of
    det(A - l B) = b00 b11 l^2 - (a00 b11 + a11 b00 - a10 b01) l + (a00 a11 - a01 a10)."""
    A = np.array(A); B = np.array(B)
    ar, ai, b = np.zeros(2), np.zeros(2), np.zeros(2)
    O.lib().orc_eig22(O.ptr(O.fmat(A)), 2, O.ptr(O.fmat(B)), 2, O.ptr(ar), O.ptr(ai), O.ptr(b))
    coef = [B[0, 0] * B[1, 1], -(A[0, 0] * B[1, 1] + A[1, 1] * B[0, 0] - A[1, 0] * B[0, 1]),
            A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0]]
    ref = np.roots(coef)
    assert _max_matched_rel_err((ar + 1j * ai) / b, ref) < 1e-13


def test_oracle_job_and_comp_variants():
    """JOB = 'E' and COMPQ/COMPZ = 'N' give the same spectrum as the full Schur form (PDHGEQZ.f:132-150);
    the factors that are computed stay orthogonal."""
    n = 90
    H, T = Q.pencil_ht_random(n, 13)
    for job, cq, cz in ((0, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)):
        S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, job=job, compq=cq, compz=cz)
        assert info == 0
        ratio, _, _ = Q.match_eigs(ar, ai, be, H, T)
        assert ratio < 1.0
        for on, M in ((cq, Q2), (cz, Z2)):
            if on:
                assert np.linalg.norm(M.T @ M - np.eye(n)) / (n * Q.EPS) < 10


def test_oracle_ilo_ihi_subblock():
    """An active block [ILO, IHI] of a pencil that is already triangular outside it: Q and Z only
    touch the block's rows and columns, the residuals hold, and the block's eigenvalues are LAPACK's."""
    n, ilo, ihi = 80, 7, 65            # 0-based bounds of the block
    H, T = Q.pencil_ht_random(n, 14)
    H[ilo, ilo - 1] = 0.0
    H[ihi + 1, ihi] = 0.0
    S, P, ar, ai, be, Q2, Z2, info, st = O.hgeqz(H, T, compq=1, compz=1, ilo=ilo, ihi=ihi)
    assert info == 0
    ra, rb, rq, rz = Q.compres(H, T, S, P, Q2, Z2)
    assert ra < 1e-14 and rb < 1e-14 and rq < 10 and rz < 10
    I = np.eye(n)
    assert np.array_equal(Q2[:ilo], I[:ilo]) and np.array_equal(Q2[ihi + 1:], I[ihi + 1:])
    assert np.array_equal(Z2[:, :ilo], I[:, :ilo]) and np.array_equal(Z2[:, ihi + 1:], I[:, ihi + 1:])
    b = slice(ilo, ihi + 1)
    ratio, _, _ = Q.match_eigs(ar[b], ai[b], be[b], H[b, b], T[b, b])
    assert ratio < 1.0


# ---- DTGEX2 (src/reorder/dtgex2.f:221-860) as the tree ships it ------------------------------------
def _orc_swap(S, T, n1, n2, n=None, j1=0):
    m = S.shape[0]
    Sf, Tf, Qf, Zf = O.fmat(S), O.fmat(T), O.fmat(np.eye(m)), O.fmat(np.eye(m))
    r = O.lib().orc_tgex2(m, O.ptr(Sf), m, O.ptr(Tf), m, O.ptr(Qf), m, O.ptr(Zf), m, j1, n1, n2)
    return r, Sf, Tf, Qf, Zf


def test_tgex2_agrees_with_lapack_dtgexc():
    """Accepted swaps of blocks cut from random generalized Schur forms: the same reordered pencil as
    LAPACK's DTGEXC up to the orthogonal freedom inside each 2x2 block (the standardisation of a
    complex pair is unique only up to that), i.e. Q^T Q_lapack and Z^T Z_lapack block-diagonal."""
    from scipy.linalg import lapack
    worst = 0.0
    for S, T, n1, n2 in Q.random_swap_corners(200, 3):
        m = n1 + n2
        r, Sf, Tf, Qf, Zf = _orc_swap(S, T, n1, n2)
        Sl, Tl, Ql, Zl, _, info = lapack.dtgexc(S, T, np.eye(m), np.eye(m), n1 + 1, 1)
        assert r == 0 and info == 0
        assert Q.schur_blocks(Sf) == Q.schur_blocks(Sl)
        for M in (Qf.T @ Ql, Zf.T @ Zl):
            E = M.copy()
            for j, k in Q.schur_blocks(Sf):
                E[j:j + k, j:j + k] = 0.0
            worst = max(worst, np.abs(E).max())
        assert np.linalg.norm(Qf @ Sf @ Zf.T - S) <= 1e-14 * np.linalg.norm(S)
        assert np.linalg.norm(Qf @ Tf @ Zf.T - T) <= 1e-14 * np.linalg.norm(T)
    assert worst < 1e-13


@pytest.mark.parametrize("n1,n2", [(2, 1), (1, 2)])
def test_tgex2_accepts_numerically_coincident_eigenvalues(n1, n2):
    """The tree's DTGEX2 does not look at DTGSY2's INFO (dtgex2.f:510-517): a swap whose Sylvester
    system is singular to working precision is accepted when the weak and strong tests pass.  Current
    LAPACK leaves on that INFO and rejects the same swaps, which is what pins this rule: for eigenvalues
    1e-9 and 1e-12 apart the oracle accepts with a backward-stable result while LAPACK rejects."""
    from scipy.linalg import lapack
    for d in (1e-12, 1e-9):
        S, T = Q.swap_corner_near(n1, n2, d)
        m = n1 + n2
        r, Sf, Tf, Qf, Zf = _orc_swap(S, T, n1, n2)
        assert r == 0
        assert np.linalg.norm(Qf @ Sf @ Zf.T - S) <= 20 * Q.EPS * np.linalg.norm(S)
        assert np.linalg.norm(Qf @ Tf @ Zf.T - T) <= 20 * Q.EPS * np.linalg.norm(T)
        assert np.abs(Sf[n2:, :n2]).max() == 0.0 and np.abs(np.tril(Tf, -1)).max() == 0.0
        info = lapack.dtgexc(S, T, np.eye(m), np.eye(m), n1 + 1, 1)[-1]
        assert info == 1


def test_tgex2_equal_1x1_eigenvalues_by_givens():
    """Two 1x1 blocks with the same eigenvalue and an off-diagonal pair (s12, t12) not proportional
    to (a, b): the Givens route of dtgex2.f:365-395 (f = 0, so the right rotation is exact) accepts
    and leaves the eigenvalue pair in place with an exact reconstruction."""
    S = np.array([[0.7, 1.3], [0.0, 0.7]]); T = np.array([[2.0, -0.4], [0.0, 2.0]])
    r, Sf, Tf, Qf, Zf = _orc_swap(S, T, 1, 1)
    assert r == 0
    assert Sf[1, 0] == 0.0 and Tf[1, 0] == 0.0
    assert abs(Sf[0, 0] / Tf[0, 0] - 0.35) < 1e-15 and abs(Sf[1, 1] / Tf[1, 1] - 0.35) < 1e-15
    assert np.linalg.norm(Qf @ Sf @ Zf.T - S) < 1e-15 and np.linalg.norm(Qf @ Tf @ Zf.T - T) < 1e-15


def test_tgex2_power_of_two_scaling_is_exact():
    """The corner is scaled by its largest entries before anything is decided (dtgex2.f:317-345), so
    scaling S by 2^k and T by 2^l scales the result bit for bit and leaves Q, Z unchanged bit for bit
    (a rule that solves the unscaled Sylvester system or uses an unscaled joint threshold does not)."""
    for S, T, n1, n2 in Q.random_swap_corners(24, 5):
        r0, S0, T0, Q0, Z0 = _orc_swap(S, T, n1, n2)
        for k, l in ((20, -30), (-40, 17)):
            r1, S1, T1, Q1, Z1 = _orc_swap(S * 2.0 ** k, T * 2.0 ** l, n1, n2)
            assert r1 == r0
            if r0 == 0:
                assert np.array_equal(S1, S0 * 2.0 ** k) and np.array_equal(T1, T0 * 2.0 ** l)
                assert np.array_equal(Q1, Q0) and np.array_equal(Z1, Z0)


def test_tgex2_rejection_leaves_the_pencil_unchanged():
    """Two 2x2 blocks 1e-3 apart with O(1) coupling: both re-triangularisations leave ||S21|| about
    2.4x above THRESH = 20 eps ||(S/MA, T/MB)||_F, so the swap is rejected (INFO = 1) and nothing
    changes (dtgex2.f:643-650, :844-848)."""
    S, T = Q.swap_corner_near(2, 2, 1e-3)
    S[2:, 2:] = np.array([[0.3 + 1e-3, 1.0], [-1e-6, 0.3 + 1e-3]])
    r, Sf, Tf, Qf, Zf = _orc_swap(S, T, 2, 2)
    assert r == 1
    assert np.array_equal(Sf, S) and np.array_equal(Tf, T) and np.array_equal(Qf, np.eye(4))


def test_oracle_tgord_matches_scipy_ordqz():
    """PDTGORD semantics (src/reorder/PDTGORD.f:1-4): the eigenvalues inside the unit disk moved to the
    top of a real generalized Schur form.  The leading m columns of Q and Z span the deflating
    subspaces that LAPACK's DTGSEN (scipy.linalg.ordqz) returns, the reordered pencil is equivalent to
    the input, and SELECT comes back as the leading m positions."""
    rng = np.random.default_rng(0)
    for n in (40, 90):
        A = rng.standard_normal((n, n)); B = rng.standard_normal((n, n))
        S, T, Qm, Zm = sl.qz(A, B, output="real")
        sel = (np.abs(Q.diag_eigs(S, T)) < 1).astype(np.int32)
        S2, T2, Q2, Z2, so, m, info = O.tgord(S, T, Qm, Zm, sel)
        assert info == 0 and m == int(sel.sum())
        assert np.array_equal(so, (np.arange(n) < m).astype(np.int32))
        ev = Q.diag_eigs(S2, T2)
        assert (np.abs(ev[:m]) < 1).all() and (np.abs(ev[m:]) >= 1).all()
        assert np.linalg.norm(Q2 @ S2 @ Z2.T - A) < 1e-13 * np.linalg.norm(A)
        assert np.linalg.norm(Q2 @ T2 @ Z2.T - B) < 1e-13 * np.linalg.norm(B)
        _, _, _, _, Qs, Zs = sl.ordqz(A, B, sort="iuc", output="real")
        for X, Y in ((Q2, Qs), (Z2, Zs)):
            assert np.abs(X[:, :m] @ X[:, :m].T - Y[:, :m] @ Y[:, :m].T).max() < 1e-12
