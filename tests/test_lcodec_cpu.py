"""Host logic of the learned snapshot codec: block ordering and the eigenvector training step."""
import numpy as np

from konglong_amd.learned import block_features, from_block_features, leading_eigvecs, FEATURES
from tests.helpers import bf16_round


def test_block_features_order_and_inverse():
    ny, nx, nz = 16, 24, 8
    v = np.arange(ny * nx * nz, dtype=np.float32).reshape(ny, nx, nz)
    X = block_features(v)
    assert X.shape == ((ny // 8) * (nx // 8) * (nz // 8), FEATURES)
    # block b = (by*nbx + bx)*nbz + bz, feature f = (ly*8 + lx)*8 + lz
    nbx, nbz = nx // 8, nz // 8
    for b, f in [(0, 0), (1, 7), (2, 8), (4, 511), (5, 77)]:
        bz = b % nbz
        bx = (b // nbz) % nbx
        by = b // (nbz * nbx)
        ly, lx, lz = f >> 6, (f >> 3) & 7, f & 7
        assert X[b, f] == v[by * 8 + ly, bx * 8 + lx, bz * 8 + lz]
    assert np.array_equal(from_block_features(X, v.shape), v)


def test_leading_eigvecs_recover_a_planted_subspace():
    """Blocks drawn from a 32-dimensional subspace: the trained W spans it, so the code is lossless
    up to rounding, and W has orthonormal columns with the deterministic sign choice."""
    rng = np.random.default_rng(3)
    B, _ = np.linalg.qr(rng.standard_normal((FEATURES, 32)))
    coeff = rng.standard_normal((4000, 32)) * np.linspace(3, 1, 32)
    X = coeff @ B.T
    W, ev = leading_eigvecs(X.T @ X, 32)
    assert np.allclose(W.T.astype(np.float64) @ W, np.eye(32), atol=1e-5)
    R = (X @ W) @ W.T
    assert np.linalg.norm(R - X) / np.linalg.norm(X) < 1e-5
    assert ev[32] / ev[0] < 1e-10
    idx = np.abs(W).argmax(axis=0)
    assert np.all(W[idx, np.arange(32)] > 0)


def test_leading_eigvecs_is_the_best_linear_code():
    """Against random orthonormal codes of the same width, the eigenvector code has the smaller
    reconstruction error on the training blocks (it is the least-squares optimum)."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((3000, FEATURES)) @ rng.standard_normal((FEATURES, FEATURES)) * 0.05
    W, _ = leading_eigvecs(X.T @ X, 32)
    err = np.linalg.norm(X - (X @ W) @ W.T)
    for s in range(5):
        Q, _ = np.linalg.qr(np.random.default_rng(100 + s).standard_normal((FEATURES, 32)))
        assert err < np.linalg.norm(X - (X @ Q) @ Q.T)


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-39, 65504.0], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    want = np.array([1.0, 1.0, 1.0078125, -2.5, 3.0e-39, 65536.0], np.float32)
    got = bf16_round(x)
    assert np.array_equal(got[[0, 1, 2, 3, 5]], want[[0, 1, 2, 3, 5]])
    assert abs(got[4] - want[4]) <= 2.0 ** -133
