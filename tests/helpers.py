"""Shared input generators and comparison helpers for the tests."""
import math

import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def smooth_field(shape, lo, hi, seed, passes=3):
    """Seeded smooth random field in [lo, hi] (box-filtered noise)."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape)
    for _ in range(passes):
        for ax in range(3):
            f = (np.roll(f, 1, ax) + f + np.roll(f, -1, ax)) / 3.0
    f = (f - f.min()) / (f.max() - f.min() + 1e-30)
    return (lo + (hi - lo) * f).astype(np.float32)


def bandlimited_field(shape, seed, cut=0.3):
    """A random (ny, nx, nz) field with no energy beyond `cut` of Nyquist on any axis and none at k = 0."""
    rng = np.random.default_rng(seed)
    spec = np.fft.rfftn(rng.standard_normal(shape))
    f = [np.abs(np.fft.fftfreq(n)) for n in shape[:2]] + [np.abs(np.fft.rfftfreq(shape[2]))]
    mask = (f[0][:, None, None] < cut / 2) & (f[1][None, :, None] < cut / 2) & (f[2][None, None, :] < cut / 2)
    spec = np.where(mask, spec, 0.0)
    spec[0, 0, 0] = 0.0
    return np.fft.irfftn(spec, s=shape, axes=(0, 1, 2))


def gaussian_blob(ny, nx, nz, cy, cx, cz, sigma):
    y, x, z = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    r2 = (y - cy) ** 2 + (x - cx) ** 2 + (z - cz) ** 2
    return np.exp(-r2 / (2.0 * sigma * sigma)).astype(np.float32)


def tilt_axis(shape, seed):
    """pjx, pjy from smooth tilt (0..45 deg) and azimuth (0..360 deg), bgprtm/ModelLoader.cpp:303-321."""
    th = smooth_field(shape, 0.0, 45.0, seed) * np.float32(np.pi / 180)
    ph = smooth_field(shape, 0.0, 360.0, seed + 1) * np.float32(np.pi / 180)
    st = np.sin(th)
    return (-st * np.cos(ph)).astype(np.float32), (-st * np.sin(ph)).astype(np.float32)


def build_api_check(out_path):
    """Compile tests/cpp/api_check.cpp against the built libraries (the bgprtm drop-in headers)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
           os.path.join(root, "tests", "cpp", "api_check.cpp"),
           "-L", os.path.join(root, "konglong_amd"), "-lkonglong_rtm", "-L", os.path.join(root, "oracle"), "-loracle",
           "-Wl,-rpath," + os.path.join(root, "konglong_amd"), "-Wl,-rpath," + os.path.join(root, "oracle"),
           "-o", str(out_path)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return str(out_path)



def bf16_round(a):
    """float32 -> nearest BF16 (ties to even), returned as float32 (finite inputs)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits_to_float(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def dense_cfd(n, d, sgn):
    """Assemble B (9-band) and A (11-band) from the literals and folding rules, dense."""
    ca5 = np.array([0.861085998189822, 0.753782983214427, 0.257072513289755, 0.031326722562826, 0.000410456919800])
    cb5 = np.array([1.0, 0.753205475977546, 0.312203852849021, 0.063580085296350, 0.004597763362256])
    fold = -1.0 if sgn == 0 else 1.0  # gradient: odd extension, divergence: even
    B = np.zeros((n, n))
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(-4, 5):
            c = cb5[abs(k)]
            j = i + k
            if j < 0:
                B[i, -j] += fold * c
            elif j > n - 1:
                B[i, 2 * (n - 1) - j] += fold * c
            else:
                B[i, j] += c
    for i in range(n):
        for k in range(-5, 6):
            c = 0.0 if k == 0 else math.copysign(ca5[abs(k) - 1], k) / (2 * d)
            j = i + k
            if 5 <= i < n - 5:
                A[i, j] += c
            elif j < 0:
                A[i, -j] += -fold * c
            elif j > n - 1:
                A[i, 2 * (n - 1) - j] += -fold * c
            else:
                A[i, j] += c
    return B, A
