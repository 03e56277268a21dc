"""The int8 CRT product (k_crt_gemm.hip) against exact and FP64 references.

C = alpha op(A) op(B) + beta C is formed from the exact product of the operands
scaled per row / column to q-bit integers (q = crt_bits(nmod)); the only error is
that scaling cut, |C - A B|_ij <= 2^(3 - q) sqrt(K) |A_i.| |B_.j| (plus FP64
rounding of the result).  Integer operands small enough to survive the cut
multiply exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191,
          181, 179, 173]
Q_BITS = {s: int(np.floor((np.log2(float(np.prod(np.array(MODULI[:s], dtype=object)))) - 1) / 2))
          for s in range(6, 21)}


def _diag():
    from uncertaintyte_amd import diag
    return diag


def _bound(A, B, q, extra=None):
    K = A.shape[1]
    ra = np.linalg.norm(A, axis=1)
    cb = np.linalg.norm(B, axis=0)
    b = 2.0 ** (3 - q) * np.sqrt(K) * np.outer(ra, cb)
    exact = A @ B
    return b + 4 * np.finfo(float).eps * np.abs(exact) + (0 if extra is None else extra)


@pytest.mark.parametrize("nmod", [6, 8, 10, 12, 13, 16, 20])
def test_crt_integer_operands_exact(nmod):
    """Integer operands whose row / column norms stay below 2^(q-3) are not cut:
    the product is the exact integer product, bit for bit (6, 8, 10, 12, 13, 16
    and 20 moduli rebuild on 3, 4, 5, 6, 7, 8 and 10 16-bit limbs; from 16
    moduli the cut integers exceed 2^52 and the residues take two steps)."""
    d = _diag()
    rng = np.random.default_rng(nmod)
    M, N, K = 300, 260, 700
    lim = 1 if nmod == 6 else 1000
    A = rng.integers(-lim, lim + 1, (M, K)).astype(np.float64)
    B = rng.integers(-lim, lim + 1, (K, N)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    C, _ = d.gemm_ex(A, B, nmod=nmod)
    assert np.array_equal(C, exact)


@pytest.mark.parametrize("nmod", [6, 10, 13])
def test_crt_error_bound_random(nmod):
    d = _diag()
    rng = np.random.default_rng(100 + nmod)
    M, N, K = 517, 333, 1029
    # rows of very different scales: the cut is relative to each row / column
    A = rng.standard_normal((M, K)) * np.exp2(rng.integers(-30, 30, (M, 1)))
    B = rng.standard_normal((K, N)) * np.exp2(rng.integers(-30, 30, (1, N)))
    C, _ = d.gemm_ex(A, B, nmod=nmod)
    err = np.abs(C - A @ B)
    bnd = _bound(A, B, Q_BITS[nmod])
    assert np.all(err <= bnd), float(np.max(err / bnd))
    # and the cut is not vacuous: the error is within a factor 2^12 of the bound's scale
    rel = np.max(err / (np.outer(np.linalg.norm(A, axis=1), np.linalg.norm(B, axis=0))))
    assert rel > 2.0 ** (-Q_BITS[nmod] - 16)


CASES = [
    # (M, N, K, trans_b, tri, beta)
    (700, 640, 650, True, 4, 0.0),    # W = A21 X11^T, X11 lower (GEMM_BT_LOWER)
    (650, 650, 700, True, 8, 1.0),    # A22 -= A21 A21^T, lower C only (GEMM_C_LOWER)
    (700, 650, 650, False, 2, 0.0),   # W = A21 X11, X11 lower (GEMM_B_LOWER)
    (700, 650, 700, False, 1, 0.0),   # X21 = -X22 W, X22 lower (GEMM_A_LOWER)
    (64, 64, 64, False, 0, 0.5),
    (1, 300, 1, True, 0, 0.0),
    (1100, 1300, 1500, False, 0, -2.0),
]


@pytest.mark.parametrize("M,N,K,trans_b,tri,beta", CASES)
def test_crt_operand_flags(M, N, K, trans_b, tri, beta):
    """Flags as in the Z+ recursion: masked parts of the operands hold garbage that
    must be ignored; with GEMM_C_LOWER the strict upper triangle of C is untouched."""
    d = _diag()
    rng = np.random.default_rng(M + 3 * N + 7 * K + tri)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K) if trans_b else (K, N))
    C0 = rng.standard_normal((M, N))
    Am, Bm = A.copy(), B.copy()
    if tri & 1:
        Am = np.tril(A)
        A = Am + np.triu(np.full_like(A, 1e30), 1)
    if tri & 2:
        Bm = np.tril(B)
        B = Bm + np.triu(np.full_like(B, -1e30), 1)
    if tri & 4:
        Bm = np.tril(B)
        B = Bm + np.triu(np.full_like(B, 3e29), 1)
    opB = Bm.T if trans_b else Bm
    alpha = -1.0 if tri & 8 else 1.0
    ref = alpha * (Am @ opB) + beta * C0
    C, _ = d.gemm_ex(A, B, C0, alpha=alpha, beta=beta, trans_b=trans_b, tri=tri, nmod=10)
    bnd = _bound(Am, opB, Q_BITS[10], 4 * np.finfo(float).eps * np.abs(beta * C0))
    if tri & 8:
        low = np.tril(np.ones((M, N), bool))
        assert np.array_equal(C[~low], C0[~low])
        assert np.all(np.abs(C - ref)[low] <= bnd[low])
    else:
        assert np.all(np.abs(C - ref) <= bnd), float(np.max(np.abs(C - ref) / bnd))


def test_crt_run_to_run_identical():
    d = _diag()
    rng = np.random.default_rng(5)
    A = rng.standard_normal((900, 800))
    B = rng.standard_normal((800, 700))
    C1, _ = d.gemm_ex(A, B, nmod=10)
    C2, _ = d.gemm_ex(A, B, nmod=10)
    assert np.array_equal(C1, C2)


def test_crt_rejects_bad_moduli_count():
    d = _diag()
    from uncertaintyte_amd import UncError
    A = np.ones((4, 4))
    for nmod in (5, 21):
        with pytest.raises(UncError):
            d.gemm_ex(A, A, nmod=nmod)
