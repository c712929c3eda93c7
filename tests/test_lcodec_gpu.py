"""GPU parity of the learned snapshot codec's matrix-core kernels against numpy.

Encode: z = bf16(x) @ bf16(W) accumulated in FP32, stored as BF16 -- compared with the FP64 product
rounded to BF16 (at most one BF16 ulp apart where the FP32 sum rounds across a BF16 boundary); the
int8 residual against x^ = bf16(z) @ bf16(W)^T with scale max|r|/127 (codes within one step of the
FP64 quantisation, rounding ties aside).  Decode: x^ + q*scale (rel 1e-6 against numpy).
Gram: sum bf16(s x)^T bf16(s x) (rel 1e-6).  Fused decode + imaging: equal, bit for bit, to decode
followed by the FP32 imaging kernel."""
import numpy as np
import pytest

from konglong_amd import _lib
from konglong_amd._lib import DeviceArray, check, lib
from konglong_amd.learned import LearnedCompressor, GramAccumulator, block_features, from_block_features
from tests.helpers import bf16_round, bf16_bits_to_float, rel_l2, smooth_field

pytestmark = pytest.mark.gpu

SHAPE = (24, 40, 48)   # ny, nx, nz: 3*5*6 = 90 blocks (two full 32-block tiles and a partial one)


def _field(seed):
    rng = np.random.default_rng(seed)
    return (smooth_field(SHAPE, -1.0, 1.0, seed) + np.float32(0.1) * rng.standard_normal(SHAPE)).astype(np.float32)


def _orthonormal(nl, seed):
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((512, nl)))
    return Q.astype(np.float32)


@pytest.mark.parametrize("nl", [32, 64])
def test_encode_decode_match_numpy(klrtm, nl):
    print(f"[lcodec] encode/decode nl={nl} shape={SHAPE}", flush=True)
    W = _orthonormal(nl, nl)
    c = LearnedCompressor(W)
    x = _field(1)
    dx = DeviceArray.from_host(x)
    code = c.new_code(SHAPE)
    c.encode(dx, code)
    zb, sc, q = c.split_code(code.download(), SHAPE)
    z = bf16_bits_to_float(zb)
    X = block_features(x).astype(np.float64)
    Wb = bf16_round(W).astype(np.float64)
    zref = bf16_round((bf16_round(X.astype(np.float32)).astype(np.float64) @ Wb).astype(np.float32))
    ulp = np.abs(z - zref) / np.maximum(np.abs(zref), 1e-30)
    assert rel_l2(z, zref) < 1e-3 and np.mean(z != zref) < 0.02 and ulp.max() <= 2.0 ** -7, (rel_l2(z, zref), ulp.max())
    # residual against the GPU's own code
    xh = z.astype(np.float64) @ Wb.T
    r = X - xh
    m = np.abs(r).max(axis=1)
    assert np.allclose(sc, m / 127.0, rtol=1e-5, atol=0)
    qref = np.clip(np.rint(r * (127.0 / m)[:, None]), -127, 127)
    assert np.abs(q - qref).max() <= 1 and np.mean(q != qref) < 1e-3
    # decode
    out = DeviceArray(SHAPE)
    c.decode(code, out)
    xr = out.download()
    xref = from_block_features(xh + q.astype(np.float64) * sc.astype(np.float64)[:, None], SHAPE)
    assert rel_l2(xr, xref) < 1e-6
    # the residual makes the code accurate whatever W is: error <= half a quantisation step
    assert np.abs(block_features(xr) - X).max() <= 0.5 * sc.max() * 1.01 + 1e-6
    c.close()


def test_gram_matches_numpy(klrtm):
    x = _field(2)
    s = np.float32(0.7)
    acc = GramAccumulator()
    dx = DeviceArray.from_host(x)
    acc.add(dx.ptr, SHAPE, s)
    acc.add(dx.ptr, SHAPE, s)        # accumulates across calls
    G = acc.matrix()
    Xb = bf16_round(block_features(x) * s).astype(np.float64)
    Gref = 2.0 * (Xb.T @ Xb)
    assert rel_l2(G, Gref) < 1e-6
    assert np.array_equal(G, G.T)


def test_fused_imaging_equals_decode_then_correlate(klrtm):
    nl = 32
    W = _orthonormal(nl, 9)
    c = LearnedCompressor(W)
    src, rec = _field(3), _field(4)
    dsrc, drec = DeviceArray.from_host(src), DeviceArray.from_host(rec)
    code = c.new_code(SHAPE)
    c.encode(dsrc, code)
    dec = DeviceArray(SHAPE)
    c.decode(code, dec)
    ny, nx, nz = SHAPE
    outs = {}
    for name in ("ref", "fused"):
        img, ill = DeviceArray(SHAPE), DeviceArray(SHAPE)
        img64, ill64 = DeviceArray(SHAPE, np.float64), DeviceArray(SHAPE, np.float64)
        if name == "ref":
            check(lib().klrtm_ic_correlate(dec.ptr, drec.ptr, nx, ny, nz, 0.8, None, 0, img.ptr, ill.ptr, img64.ptr,
                                           ill64.ptr))
        else:
            check(lib().klrtm_lcodec_correlate(c.h, code.ptr, drec.ptr, nx, ny, nz, 0.8, None, 0, img.ptr, ill.ptr,
                                               img64.ptr, ill64.ptr))
        outs[name] = [a.download() for a in (img, ill, img64, ill64)]
    for a, b in zip(outs["ref"], outs["fused"]):
        assert np.array_equal(a, b)
    c.close()


def test_rejects_grids_not_multiple_of_8(klrtm):
    c = LearnedCompressor(_orthonormal(32, 1))
    v = DeviceArray((8, 8, 12))
    code = DeviceArray((4096,), np.uint8)
    assert lib().klrtm_lcodec_encode(c.h, v.ptr, code.ptr, 8, 8, 12) != 0
    assert b"multiples of 8" in lib().klrtm_last_error()
