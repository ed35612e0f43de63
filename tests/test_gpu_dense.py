"""GPU: the batched dense kernels (MFMA complex GEMM, partial-pivot LU solve, CORTH+COMQR2 eigensolver)
against numpy, through the C-ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_zgemm_shapes(ctx):
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(1)
    shapes = [(1, 1, 1), (16, 16, 4), (17, 33, 5), (32, 32, 32), (121, 121, 121), (70, 3, 121), (121, 1, 121),
              (5, 64, 2), (241, 241, 241)]
    As = [crand(rng, m, k) for m, n, k in shapes]
    Bs = [crand(rng, k, n) for m, n, k in shapes]
    Cs = rcwa.zgemm_batched(As, Bs, ctx)
    for A, B, C in zip(As, Bs, Cs):
        ref = A @ B
        assert np.max(np.abs(C - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref))) * A.shape[1]


def test_zgesv(ctx):
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(2)
    sizes = [(1, 1), (2, 3), (16, 16), (17, 1), (41, 41), (121, 121), (121, 2), (250, 7)]
    As = [crand(rng, n, n) for n, r in sizes]
    Bs = [crand(rng, n, r) for n, r in sizes]
    out = rcwa.zgesv_batched(As, Bs, ctx)
    for A, B, (X, info) in zip(As, Bs, out):
        assert info == 0
        berr = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X) + np.linalg.norm(B))
        assert berr <= 1e-14


def test_zgesv_zero_pivot(ctx):
    from stablercwa_amd import rcwa
    A = np.zeros((5, 5), dtype=complex)
    A[0, 0] = 1.0
    (X, info), = rcwa.zgesv_batched([A], [np.ones((5, 1), dtype=complex)], ctx)
    assert info == 1


def _check_eig(A, w, V, info, tol=1e-13):
    assert info == 0
    n = A.shape[0]
    anorm = np.linalg.norm(A)
    res = np.max(np.linalg.norm(A @ V - V * w[None, :], axis=0) / (anorm * np.linalg.norm(V, axis=0)))
    assert res <= tol, res
    assert np.allclose(np.linalg.norm(V, axis=0), 1.0, atol=1e-13)
    ref = np.linalg.eigvals(A)
    # every computed eigenvalue is close to a reference eigenvalue
    d = np.min(np.abs(w[:, None] - ref[None, :]), axis=1)
    assert np.max(d) <= 1e-9 * max(1.0, anorm), np.max(d)
    # eigenvectors span the space (no collapsed pairs)
    assert np.linalg.cond(V) < 1e10
    return res


def test_zgeev_matrices(ctx):
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(3)
    mats = [crand(rng, n, n) for n in (1, 2, 3, 7, 41, 64, 121)]
    H = crand(rng, 50, 50); mats.append(H + H.conj().T)              # Hermitian
    D = np.diag(np.repeat(np.arange(1, 11), 2).astype(complex)); mats.append(D + 0.0)   # exactly repeated
    # RCWA-like Toeplitz layer matrix: Kx^2 - [[eps]]
    n = 61
    c = crand(rng, 2 * n - 1) * 0.2
    T = np.array([[c[i - j + n - 1] for j in range(n)] for i in range(n)])
    kx = 0.9 - (np.arange(n) - n // 2) * 1.8
    mats.append(np.diag(kx ** 2) - T)
    out = rcwa.zgeev_batched(mats, ctx)
    for A, (w, V, info) in zip(mats, out):
        _check_eig(A, w, V, info)


def test_zgeev_residual_not_worse_than_reference(ctx, ref_lib):
    """The eigen-pair residual max_j ||A v_j - l_j v_j|| / (||A||_F ||v_j||) (eig_residual, reference
    eig.hpp:38-69) over 80 random matrices (orders 41, 121, 241) and 32 of the reference's own planar
    layer matrices (mirror grating, M = 60, TE and TM): the GPU's residuals are not worse than the
    reference QR's on the same matrices (conftest.not_worse_than), and none exceeds 2e-15."""
    import ctypes
    from oracle import ref
    from stablercwa_amd import rcwa
    from stablercwa_amd.workloads import mirror_stack
    from conftest import not_worse_than
    rng = np.random.default_rng(123)
    groups = [[crand(rng, n, n) for _ in range(k)] for n, k in ((41, 32), (121, 32), (241, 16))]
    f = mirror_stack().flat()
    layer = []
    for pol in (0, 1):
        for lam in np.linspace(0.95, 1.15, 16):
            A = np.zeros(2 * 121 * 121)
            assert ref_lib.ref_layer_matrix(ctypes.c_void_p(f.ctypes.data), 0, ctypes.c_double(lam),
                                            ctypes.c_double(66.36), 60, pol, ctypes.c_void_p(A.ctypes.data)) == 0
            layer.append(A.view(np.complex128).reshape((121, 121), order="F").copy())
    groups.append(layer)
    gpu, refr = [], []
    for mats in groups:
        for A, (w, V, info) in zip(mats, rcwa.zgeev_batched(mats, ctx)):
            gpu.append(_check_eig(A, w, V, info, tol=2e-15))
            _, _, rinfo, rres = ref.eig(ref_lib, A)
            assert rinfo == 0
            refr.append(rres)
    not_worse_than(gpu, refr, "eig residual")


def test_zgesv_every_panel_path(ctx):
    """Orders on both sides of each LU panel form (one wave: up to 64 and 128 rows; the workgroup: up to 256
    and 512 rows; global memory above) and of the last, narrower panel, with right-hand-side counts on
    both sides of the MFMA / VALU substitution update; entries of equal modulus force pivot ties."""
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(21)
    cases = [(63, 1), (64, 64), (65, 8), (127, 7), (128, 9), (129, 129), (144, 3), (255, 1), (256, 256),
             (257, 2), (511, 16), (512, 5), (513, 513), (600, 1)]
    As, Bs = [], []
    for n, r in cases:
        A = crand(rng, n, n)
        if n in (65, 129, 257):
            A = (rng.integers(0, 2, (n, n)) * 2 - 1) + 1j * (rng.integers(0, 2, (n, n)) * 2 - 1)  # ties
            A = A + 0.05 * np.eye(n)
        As.append(A)
        Bs.append(crand(rng, n, r))
    out = rcwa.zgesv_batched(As, Bs, ctx)
    for (n, r), A, B, (X, info) in zip(cases, As, Bs, out):
        assert info == 0, n
        berr = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X) + np.linalg.norm(B))
        assert berr <= 1e-14, (n, r, berr)


def test_zgeev_split_windows(ctx):
    """Matrices whose QR iteration splits into separate windows (weakly coupled diagonal blocks), at orders
    of each eigen-kernel instantiation: the queued off-window updates must land on the right rows and
    columns whatever order the windows come in."""
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(22)
    mats = []
    for n in (90, 200, 300, 530):
        A = crand(rng, n, n)
        cut = [0, n // 3, 2 * n // 3, n]
        for a in range(3):
            for b in range(3):
                if a != b:
                    A[cut[a]:cut[a + 1], cut[b]:cut[b + 1]] *= 1e-13
        mats.append(A)
    for A in mats:      # one call each: a batch runs the instantiation of its largest order
        (w, V, info), = rcwa.zgeev_batched([A], ctx)
        _check_eig(A, w, V, info, tol=1e-12)


def test_zgeev_result_independent_of_batch(ctx):
    """A batch runs the eigen instantiation of its largest order (one wave carries the QR rotation chain
    up to order 256, the whole workgroup above); both compute the same rotations, so a matrix's
    eigenpairs are bitwise the same alone and next to a larger matrix."""
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(23)
    A, B = crand(rng, 150, 150), crand(rng, 300, 300)
    (w1, V1, i1), = rcwa.zgeev_batched([A], ctx)
    (w2, V2, i2), _ = rcwa.zgeev_batched([A, B], ctx)
    assert i1 == 0 and i2 == 0
    assert np.array_equal(w1, w2)
    assert np.array_equal(V1, V2)
