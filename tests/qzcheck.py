"""Shared checks for the QZ tests: pencil generators, PCOMPRES-style residuals and the
condition-weighted eigenvalue comparison used to pin results (the reference's tests only check
residuals, testing/EXRAND1.f:329-332; eigenvalue sets are compared here against LAPACK DGGES)."""
import numpy as np
import scipy.linalg as sl
from scipy.optimize import linear_sum_assignment

EPS = np.finfo(np.float64).eps


# ---- generators (SplitMix64 + Box-Muller, identical to the oracle's orc_fill_normal) ---------
def normal(seed, shape):
    import oracle_ffi as O
    return O.normal(seed, shape)


def uniform(seed, shape):
    import oracle_ffi as O
    return O.uniform(seed, shape)


def pencil_randn(n, seed):
    """configs 1/2: A, B i.i.d. N(0,1)."""
    return normal(2 * seed + 1, (n, n)), normal(2 * seed + 2, (n, n))


def pencil_stress(n, seed, zero_frac=0.05, nnz_row=10, grade=3.0):
    """config 3: A = sparse(nnz_row N(0,1) per row at uniform columns) + 0.5 I, B = triu(N(0,1))
    with zero_frac of its diagonal set to 0, rows/columns graded by 10^U(-grade, grade)."""
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        cols = rng.choice(n, size=min(nnz_row, n), replace=False)
        A[i, cols] = rng.standard_normal(len(cols))
    A += 0.5 * np.eye(n)
    B = np.triu(rng.standard_normal((n, n)))
    zi = rng.choice(n, size=max(1, int(round(zero_frac * n))), replace=False)
    B[zi, zi] = 0.0
    dr = 10.0 ** rng.uniform(-grade, grade, n)
    dc = 10.0 ** rng.uniform(-grade, grade, n)
    return (dr[:, None] * A) * dc[None, :], (dr[:, None] * B) * dc[None, :]


def pencil_known(n, seed, n_inf=0, offdiag=0.5):
    """Constructed pencil with a known spectrum: A = U RA V^T, B = U RB V^T, RA, RB upper
    triangular with a mild nilpotent part; n_inf rows of RB are zero, giving exactly n_inf
    semisimple infinite eigenvalues.  Returns (A, B, finite_eigs, n_inf)."""
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sc = offdiag / np.sqrt(n)
    da = rng.uniform(1, 2, n) * np.sign(rng.standard_normal(n))
    db = rng.uniform(0.5, 2, n)
    RA = sc * np.triu(rng.standard_normal((n, n)), 1) + np.diag(da)
    RB = sc * np.triu(rng.standard_normal((n, n)), 1) + np.diag(db)
    zi = rng.choice(n, size=n_inf, replace=False) if n_inf else np.array([], dtype=int)
    RB[zi, :] = 0.0
    fin = np.setdiff1d(np.arange(n), zi)
    return U @ RA @ V.T, U @ RB @ V.T, da[fin] / db[fin], n_inf


def pencil_ht_random(n, seed):
    """Hessenberg-triangular pencil with the distribution of an HT-reduced Gaussian pencil (the
    construction of testing/EXRAND1.f:243-279: N(0,1) upper parts, subdiagonal of H and diagonal
    of T drawn as chi variables)."""
    rng = np.random.default_rng(seed)
    H = np.triu(rng.standard_normal((n, n)))
    T = np.triu(rng.standard_normal((n, n)))
    for j in range(n - 1):
        H[j + 1, j] = np.sqrt(rng.chisquare(n - j - 1))
    for j in range(n):
        T[j, j] = np.sqrt(rng.chisquare(n - j))
    return H, T


# ---- checks ------------------------------------------------------------------------------------
def compres(A, B, S, T, Q, Z):
    """PCOMPRES (tools/PQZHLPS.f:19-319): resA, resB, resQ/(n eps), resZ/(n eps)."""
    n = A.shape[0]
    ra = np.linalg.norm(Q.T @ A @ Z - S) / np.linalg.norm(A)
    rb = np.linalg.norm(Q.T @ B @ Z - T) / np.linalg.norm(B)
    rq = np.linalg.norm(Q.T @ Q - np.eye(n)) / (n * EPS)
    rz = np.linalg.norm(Z.T @ Z - np.eye(n)) / (n * EPS)
    return ra, rb, rq, rz


def schur_structure(S, T):
    """(max |S below subdiag|, max |T below diag|, 2x2 blocks standardised, beta >= 0)."""
    n = S.shape[0]
    below = np.abs(np.tril(S, -2)).max() if n > 2 else 0.0
    tb = np.abs(np.tril(T, -1)).max() if n > 1 else 0.0
    ok2 = True
    j = 0
    while j < n:
        if j + 1 < n and S[j + 1, j] != 0.0:
            if j + 2 < n and S[j + 2, j + 1] != 0.0:
                ok2 = False          # two consecutive non-zero subdiagonals: not quasi-triangular
            if T[j, j + 1] != 0.0 or T[j, j] < 0 or T[j + 1, j + 1] < 0:
                ok2 = False
            j += 2
        else:
            j += 1
    return below, tb, ok2


def chordal(a1, b1, a2, b2):
    """chordal distance between homogeneous eigenvalues (a1:b1) and (a2:b2) (complex a)."""
    num = np.abs(a1 * b2 - a2 * b1)
    den = np.sqrt(np.abs(a1) ** 2 + np.abs(b1) ** 2) * np.sqrt(np.abs(a2) ** 2 + np.abs(b2) ** 2)
    return num / np.maximum(den, 1e-300)


def lapack_eigs(A, B):
    """LAPACK (scipy) homogeneous eigenvalues and their condition numbers
    kappa_j = ||x|| ||y|| / sqrt(|y^H A x|^2 + |y^H B x|^2) (Stewart-Sun / DTGSNA style)."""
    w, vl, vr = sl.eig(A, B, left=True, right=True, homogeneous_eigvals=True)
    a, b = w
    cond = np.empty(len(a))
    for j in range(len(a)):
        x, y = vr[:, j], vl[:, j]
        den = np.sqrt(np.abs(y.conj() @ A @ x) ** 2 + np.abs(y.conj() @ B @ x) ** 2)
        cond[j] = np.linalg.norm(x) * np.linalg.norm(y) / max(den, 1e-300)
    return a, b, cond


def match_eigs(ar, ai, be, A, B, fudge=100.0):
    """Condition-weighted comparison of (ar + i ai, be) with LAPACK's eigenvalues of (A, B).
    Returns (worst ratio chordal / (fudge n eps ||(A,B)|| kappa), max chordal, #inf ours)."""
    n = A.shape[0]
    a1 = np.asarray(ar) + 1j * np.asarray(ai)
    b1 = np.asarray(be).astype(complex)
    a2, b2, cond = lapack_eigs(A, B)
    D = chordal(a1[:, None], b1[:, None], a2[None, :], b2[None, :])
    ri, ci = linear_sum_assignment(D)
    nab = np.sqrt(np.linalg.norm(A) ** 2 + np.linalg.norm(B) ** 2)
    tol = fudge * n * EPS * nab * cond[ci]
    ratio = D[ri, ci] / np.maximum(tol, 1e-300)
    return float(ratio.max()), float(D[ri, ci].max()), int((np.asarray(be) == 0).sum())


def match_known(ar, ai, be, known, A, B, fudge=100.0):
    """Finite computed eigenvalues against a known finite spectrum, with the same condition-weighted
    chordal tolerance as match_eigs (condition numbers from LAPACK's eigenvectors of (A, B), matched
    to the known values).  Returns the worst ratio distance / tolerance."""
    n = A.shape[0]
    be = np.asarray(be)
    fin = be != 0
    ev = (np.asarray(ar) + 1j * np.asarray(ai))[fin] / be[fin]
    a2, b2, cond = lapack_eigs(A, B)
    lf = np.abs(b2) > 0
    lam = a2[lf] / b2[lf]
    kap = cond[lf]
    # condition number of each known eigenvalue: that of the nearest LAPACK eigenvalue
    D0 = np.abs(known[:, None] - lam[None, :])
    kk = kap[np.argmin(D0, axis=1)]
    D = chordal(ev[:, None], np.ones((len(ev), 1)), known[None, :].astype(complex), np.ones((1, len(known))))
    ri, ci = linear_sum_assignment(D)
    nab = np.sqrt(np.linalg.norm(A) ** 2 + np.linalg.norm(B) ** 2)
    tol = fudge * n * EPS * nab * kk[ci]
    return float((D[ri, ci] / tol).max())


# ---- adjacent-block swap cases (DTGEX2, src/reorder/dtgex2.f) ------------------------------------
def swap_corner_near(n1, n2, delta, coup=1.0, seed=0):
    """(n1 + n2)-order corner with blocks whose eigenvalues are delta apart: a 2x2 block
    [[mu, 1], [-nu^2, mu]] (T = I) has eigenvalues mu +- i nu; the upper block sits at mu = 0.3, the
    lower one at 0.3 + delta (or, for two 2x2 blocks, with nu = delta above and nu = 1e-3 below);
    the off-diagonal blocks of S and T are coup * N(0, 1)."""
    rng = np.random.default_rng(seed)
    m = n1 + n2
    S = np.zeros((m, m)); T = np.zeros((m, m))
    blk = lambda mu, nu: (np.array([[mu, 1.0], [-nu * nu, mu]]), np.eye(2))
    if n1 == 2:
        S[:2, :2], T[:2, :2] = blk(0.3, delta)
    else:
        S[0, 0], T[0, 0] = 0.3, 1.0
    if n2 == 2:
        S[n1:, n1:], T[n1:, n1:] = blk(0.3 + (delta if n1 == 1 else 0.0), 1e-3 if n1 == 2 else delta)
    else:
        S[n1, n1], T[n1, n1] = 0.3 + (delta if n1 == 1 else 0.0), 1.0
    S[:n1, n1:] = coup * rng.standard_normal((n1, n2))
    T[:n1, n1:] = coup * rng.standard_normal((n1, n2))
    return S, T


def schur_blocks(S):
    """[(start, order)] of the diagonal blocks of a quasi-triangular S."""
    n = S.shape[0]
    out, j = [], 0
    while j < n:
        if j + 1 < n and S[j + 1, j] != 0.0:
            out.append((j, 2)); j += 2
        else:
            out.append((j, 1)); j += 1
    return out


def random_swap_corners(count, seed, n=8):
    """count adjacent-block corners (S, T, n1, n2) cut from the real generalized Schur forms of
    random N(0, 1) pencils (LAPACK DGGES via scipy), cycling over the four block-size pairs."""
    rng = np.random.default_rng(seed)
    want = [(1, 1), (1, 2), (2, 1), (2, 2)]
    out = []
    while len(out) < count:
        A = rng.standard_normal((n, n)); B = rng.standard_normal((n, n))
        S, T, _, _ = sl.qz(A, B, output="real")
        bl = schur_blocks(S)
        for k in range(len(bl) - 1):
            (j1, n1), (_, n2) = bl[k], bl[k + 1]
            if (n1, n2) == want[len(out) % 4]:
                m = n1 + n2
                out.append((S[j1:j1 + m, j1:j1 + m].copy(), T[j1:j1 + m, j1:j1 + m].copy(), n1, n2))
                break
    return out


def match_two(ar1, ai1, be1, ar2, ai2, be2, A, B, fudge=100.0):
    """Worst condition-weighted chordal distance between two computed eigenvalue sets of (A, B)
    (e.g. the GPU's and the oracle's): each eigenvalue of set 1 takes the condition number of the
    LAPACK eigenvalue it is closest to, the sets are matched by an optimal assignment, and the
    distance is divided by fudge * n * eps * ||(A, B)||_F * kappa."""
    n = A.shape[0]
    a1 = np.asarray(ar1) + 1j * np.asarray(ai1); b1 = np.asarray(be1).astype(complex)
    a2 = np.asarray(ar2) + 1j * np.asarray(ai2); b2 = np.asarray(be2).astype(complex)
    la, lb, cond = lapack_eigs(A, B)
    D0 = chordal(a1[:, None], b1[:, None], la[None, :], lb[None, :])
    r0, c0 = linear_sum_assignment(D0)
    kap = np.empty(n); kap[r0] = cond[c0]
    D = chordal(a1[:, None], b1[:, None], a2[None, :], b2[None, :])
    ri, ci = linear_sum_assignment(D)
    nab = np.sqrt(np.linalg.norm(A) ** 2 + np.linalg.norm(B) ** 2)
    ratio = D[ri, ci] / np.maximum(fudge * n * EPS * nab * kap[ri], 1e-300)
    return float(ratio.max())


def diag_eigs(S, T):
    """Eigenvalues lambda_j of a real generalized Schur form in diagonal order (complex pairs from
    their 2x2 blocks); inf where T(j, j) = 0 for a 1x1 block."""
    n = S.shape[0]
    out, j = [], 0
    while j < n:
        if j + 1 < n and S[j + 1, j] != 0.0:
            out += list(sl.eigvals(S[j:j + 2, j:j + 2], T[j:j + 2, j:j + 2])); j += 2
        else:
            out.append(S[j, j] / T[j, j] if T[j, j] != 0.0 else np.inf); j += 1
    return np.array(out, dtype=complex)
