"""Pins the oracle's first-derivative engines, first-derivative PML, ScalarVTI/ScalarTTI, the VTI/TTI
operator and the shot time loop against brute force, closed forms, physics and adjointness, in FP64,
so that a dropped term, a flipped sign, a wrong coefficient index or an off-by-one time index in the
restatement fails a check.  Each check names the reference lines it pins:

  FD first derivatives        bgprtm/FdEngine_fd.cpp:388-470
  first-derivative PML        bgprtm/FdEngine.cpp:301-350, bgprtm/PML.cpp:62-70
  ScalarVTI / ScalarTTI       bgprtm/KernelCPU.cpp:176-297
  KernelCPU::VTI / TTI        bgprtm/KernelCPU.cpp:93-123
  modelling / migration loop  bgprtm/Propagator.cpp:1101-1153, 1199-1255
"""
import math

import numpy as np
import pytest

import oracle
from oracle import pipeline
from tests.helpers import dense_cfd, smooth_field

FD, CFD = 7, 4
VTI, TTI = 2, 3
# array axes of the (ny, nx, nz) volumes for the engine's axis numbers (0 z, 1 x, 2 y)
ARR_AX = {0: 2, 1: 1, 2: 0}


def _fd_matrix(n, M, d):
    """Dense central first derivative with zero halo: (D u)_i = sum_m c1_m (u_{i+m} - u_{i-m}) / d, with the
    Taylor c1 (pinned by known answers in test_oracle_cpu) and out-of-range samples dropped."""
    c1 = oracle.taylor_coef1(M)
    D = np.zeros((n, n))
    for i in range(n):
        for m in range(1, M + 1):
            if i + m < n:
                D[i, i + m] += c1[m] / d
            if i - m >= 0:
                D[i, i - m] -= c1[m] / d
    return D


def _cfd_matrix(n, d, sgn):
    """The compact derivative as a dense operator B^-1 A (assembled from the scheme's literals and the
    boundary folding, pinned against dense solves and the compiled reference plan)."""
    B, A = dense_cfd(n, d, sgn)
    return np.linalg.solve(B, A)


def _apply(Dm, u, axis):
    ax = ARR_AX[axis]
    return np.moveaxis(np.tensordot(Dm, np.moveaxis(u, ax, 0), axes=(1, 0)), 0, ax)


def _pml_coef(npml, d, dt, vmax, eps):
    """PML::init's profile in FP64: 0.75 ((npml - i)/npml)^2 ln(1/eps) / (npml d / (dt vmax)); the FP32
    profile is pinned to this closed form in test_oracle_cpu."""
    i = np.arange(npml)
    return 0.75 * ((npml - i) / npml) ** 2 * math.log(1 / eps) / (npml * d / (dt * vmax))


def _damping(shape, axis, nb, spacing, dt, vmax, eps, scaler, top=True):
    """Product of 1/(1 + g) over the damping rounds a first-derivative output sees on its first call
    (zero state).  One round with state q = 0 maps p to q' + p/(1 - g), q' = -2g p/(1 - g^2), which is
    p/(1 + g).  Own axis at full strength, the two cross axes scaled by `scaler`; g at distance i from
    the outer edge is the profile's coef[i]."""
    ny, nx, nz = shape
    n_of = {0: nz, 1: nx, 2: ny}
    w_of = {0: (nb[4], nb[5]), 1: (nb[0], nb[1]), 2: (nb[2], nb[3])}
    d_of = {0: spacing[2], 1: spacing[0], 2: spacing[1]}
    f = np.ones(shape)
    for b in (0, 1, 2):
        if b == 2 and axis == 2 and ny <= 1:
            continue
        sc = 1.0 if b == axis else scaler
        n = n_of[b]
        g = np.zeros(n)
        for side, w in enumerate(w_of[b]):
            if w <= 1 or (b == 0 and side == 0 and not top):
                continue
            coef = _pml_coef(w, d_of[b], dt, vmax, eps)
            for i in range(w):
                pos = i if side == 0 else n - 1 - i
                g[pos] = coef[i] * sc
        # faces of one axis never overlap here (widths fit inside the axis)
        shp = [1, 1, 1]
        shp[ARR_AX[b]] = n
        f = f / (1.0 + g.reshape(shp))
    return f


def _engine(shape, spacing, engine=FD, order=8, dtype=np.float64):
    ny, nx, nz = shape
    return oracle.Engine(nx, ny, nz, *spacing, engine, order, dtype)


# ---- FD first derivatives ------------------------------------------------------------------
@pytest.mark.parametrize("order", [8, 16])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fd_first_derivative_brute_force(order, axis):
    """d{x,y,z}1 with no PML equals the dense zero-halo operator; iacc=1 subtracts vel*D(u) from out."""
    shape, spacing = (13, 19, 23), (10.0, 12.0, 7.0)
    rng = np.random.default_rng(order + axis)
    e = _engine(shape, spacing, order=order)
    vel = rng.random(shape) + 0.5
    e.set_model(vel)
    e.set_boundary([0] * 6, 1e-3, 3000.0, dim_pml=3)
    u = rng.standard_normal(shape)
    d = {0: spacing[2], 1: spacing[0], 2: spacing[1]}[axis]
    want = _apply(_fd_matrix(shape[ARR_AX[axis]], order // 2, d), u, axis)
    for isign in (1, -1):
        out = np.full(shape, 7.0)
        e.deriv1(axis, u, out, 0, isign)
        assert np.allclose(out, want, rtol=0, atol=1e-12 * np.abs(want).max())
        base = rng.standard_normal(shape)
        acc = base.copy()
        e.deriv1(axis, u, acc, 1, isign)
        assert np.allclose(acc, base - vel * want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fd_first_derivative_antisymmetric_and_exact(axis):
    """With zero halo the operator is antisymmetric, <Du, v> = -<u, Dv>; and it differentiates polynomials
    of degree <= 2M exactly where the whole stencil is inside the grid."""
    shape, spacing, M = (12, 14, 17), (9.0, 11.0, 6.0), 4
    e = _engine(shape, spacing)
    e.set_model(np.ones(shape))
    e.set_boundary([0] * 6, 1e-3, 3000.0, dim_pml=3)
    rng = np.random.default_rng(axis)
    u, v = rng.standard_normal(shape), rng.standard_normal(shape)
    du, dv = np.zeros(shape), np.zeros(shape)
    e.deriv1(axis, u, du, 0, 1)
    e.deriv1(axis, v, dv, 0, 1)
    assert abs(np.vdot(du, v) + np.vdot(u, dv)) < 1e-12 * np.linalg.norm(du) * np.linalg.norm(v)
    ax = ARR_AX[axis]
    n = shape[ax]
    d = {0: spacing[2], 1: spacing[0], 2: spacing[1]}[axis]
    x = (np.arange(n) - n / 2) * d
    for p in (1, 3, 8):
        line = (x / (n * d)) ** p
        shp = [1, 1, 1]
        shp[ax] = n
        f = np.broadcast_to(line.reshape(shp), shape).copy()
        out = np.zeros(shape)
        e.deriv1(axis, f, out, 0, 1)
        want = p * (x / (n * d)) ** (p - 1) / (n * d)
        got = np.moveaxis(out, ax, 0)[M:n - M, 0, 0]
        assert np.allclose(got, want[M:n - M], rtol=1e-9, atol=1e-12)


# ---- first-derivative PML (3D, cross-axis damping) --------------------------------------------
@pytest.mark.parametrize("engine", [FD, CFD])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pml1_first_call_closed_form(engine, axis):
    """From a clean state one call damps every cell by prod 1/(1 + g): the own axis at full strength and
    the two cross axes at pml3d_scaler, each face with its own profile and the bottom faces mirrored.
    The gradient (isign +1) and divergence (isign -1) rounds and the three components of each face
    keep separate state, so a first call of any other (axis, isign) pair sees zero state too."""
    shape, spacing = (16, 18, 20), (10.0, 12.0, 8.0)
    nb = [3, 4, 5, 2, 4, 6]
    dt, vmax, eps, sc = 1e-3, 3000.0, 1e-3, 0.3
    rng = np.random.default_rng(11 + axis + engine)
    u = rng.standard_normal(shape)
    e = _engine(shape, spacing, engine=engine, order=8)
    e.set_model(np.ones(shape))
    e.set_boundary(nb, dt, vmax, eps, True, sc, True, 3)
    e0 = _engine(shape, spacing, engine=engine, order=8)
    e0.set_model(np.ones(shape))
    e0.set_boundary([0] * 6, dt, vmax, eps, True, sc, True, 3)
    # the compact engine's divergence (isign -1) uses its own plan, so the undamped output is per sign
    raw, raw_div = np.zeros(shape), np.zeros(shape)
    e0.deriv1(axis, u, raw, 0, 1)
    e0.deriv1(axis, u, raw_div, 0, -1)
    want = raw * _damping(shape, axis, nb, spacing, dt, vmax, eps, sc)
    tol = 1e-12 * np.abs(raw).max()
    for other in (0, 1, 2):  # the other axes' components and rounds are distinct state
        if other != axis:
            e.deriv1(other, u, np.zeros(shape), 0, 1)
            e.deriv1(other, u, np.zeros(shape), 0, -1)
    got = np.zeros(shape)
    e.deriv1(axis, u, got, 0, -1)       # first divergence-round call
    assert np.allclose(got, raw_div * _damping(shape, axis, nb, spacing, dt, vmax, eps, sc), rtol=0, atol=tol)
    got = np.zeros(shape)
    e.deriv1(axis, u, got, 0, 1)        # first gradient-round call
    assert np.allclose(got, want, rtol=0, atol=tol)
    # a second gradient-round call sees the state: in cells of one face only, a repeated input p
    # leaves p (1 - g) / (1 + g)^2
    got2 = np.zeros(shape)
    e.deriv1(axis, u, got2, 0, 1)
    g = 1.0 / _damping(shape, axis, nb, spacing, dt, vmax, eps, sc) - 1.0
    iy, ix, iz = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    faces = ((ix < nb[0]) | (ix >= shape[1] - nb[1])).astype(int) + ((iy < nb[2]) | (iy >= shape[0] - nb[3])) + \
        ((iz < nb[4]) | (iz >= shape[2] - nb[5]))
    one = faces == 1
    assert one.sum() > 100
    assert np.allclose(got2[one], (raw * (1 - g) / (1 + g) ** 2)[one], rtol=0, atol=tol)
    # off the faces nothing is damped
    assert np.allclose(got2[faces == 0], raw[faces == 0], rtol=0, atol=tol)


def test_pml1_without_3d_or_top_layer():
    """pml3d off leaves only the own axis; a free surface (top PML off) leaves the top z face undamped;
    scaler 0 with pml3d on equals pml3d off bit for bit (g = 0 rounds are the identity)."""
    shape, spacing, nb = (12, 14, 16), (10.0, 10.0, 10.0), [3, 3, 3, 3, 4, 4]
    dt, vmax, eps = 1e-3, 3000.0, 1e-3
    u = np.random.default_rng(2).standard_normal(shape)
    e0 = _engine(shape, spacing)
    e0.set_model(np.ones(shape))
    e0.set_boundary([0] * 6, dt, vmax, eps, True, 0.3, True, 3)
    for axis in (0, 1, 2):
        raw = np.zeros(shape)
        e0.deriv1(axis, u, raw, 0, 1)
        e = _engine(shape, spacing)
        e.set_model(np.ones(shape))
        e.set_boundary(nb, dt, vmax, eps, False, 0.3, True, 3)
        got = np.zeros(shape)
        e.deriv1(axis, u, got, 0, 1)
        want = raw * _damping(shape, axis, nb, spacing, dt, vmax, eps, 0.0)
        assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(raw).max())
        e = _engine(shape, spacing)
        e.set_model(np.ones(shape))
        e.set_boundary(nb, dt, vmax, eps, True, 0.3, False, 3)
        got = np.zeros(shape)
        e.deriv1(axis, u, got, 0, 1)
        want = raw * _damping(shape, axis, nb, spacing, dt, vmax, eps, 0.3, top=False)
        assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(raw).max())
        a, b = _engine(shape, spacing), _engine(shape, spacing)
        for x, p3 in ((a, True), (b, False)):
            x.set_model(np.ones(shape))
            x.set_boundary(nb, dt, vmax, eps, p3, 0.0, True, 3)
        for _ in range(3):
            ga, gb = np.zeros(shape), np.zeros(shape)
            a.deriv1(axis, u, ga, 0, 1)
            b.deriv1(axis, u, gb, 0, 1)
            assert np.array_equal(ga, gb)


# ---- ScalarVTI / ScalarTTI ---------------------------------------------------------------------
def _scalar(model, g, eps, dl, n=None):
    """Apply the oracle's scalar operator to gradient vectors g (N, 3) with per-point eps/delta and
    symmetry axis n (N, 2: pjx, pjy)."""
    N = g.shape[0]
    e = oracle.Engine(N, 1, 1, 10.0, 10.0, 10.0, FD, 8, np.float64)
    pj = np.zeros((N, 2)) if n is None else n
    e.set_model(np.ones((1, N, 1)), eps.reshape(1, N, 1), dl.reshape(1, N, 1), pj[:, 0].reshape(1, N, 1).copy(),
                pj[:, 1].reshape(1, N, 1).copy())
    wx, wy, wz = (np.ascontiguousarray(g[:, i]).reshape(1, N, 1).copy() for i in range(3))
    e.scalar(model, wx, wy, wz)
    return np.stack([wx.ravel(), wy.ravel(), wz.ravel()], 1)


def _unit_axes(N, rng):
    th = rng.uniform(0, np.pi / 3, N)
    ph = rng.uniform(0, 2 * np.pi, N)
    return np.stack([-np.sin(th) * np.cos(ph), -np.sin(th) * np.sin(ph)], 1)


def test_scalar_isotropic_and_elliptic_limits():
    """eps = delta = 0 is the identity; eps = delta (elliptic) scales the part of the gradient across the
    symmetry axis by 1 + 2 eps and keeps the part along it (VTI: axis z, pz untouched)."""
    rng = np.random.default_rng(4)
    N = 500
    g = rng.standard_normal((N, 3))
    z = np.zeros(N)
    n = _unit_axes(N, rng)
    for model in (VTI, TTI):
        assert np.allclose(_scalar(model, g, z, z, n), g, rtol=0, atol=1e-15)
    eps = rng.uniform(0, 0.3, N)
    out = _scalar(VTI, g, eps, eps)
    assert np.allclose(out[:, :2], g[:, :2] * (1 + 2 * eps)[:, None], rtol=1e-14)
    assert np.array_equal(out[:, 2], g[:, 2])
    out = _scalar(TTI, g, eps, eps, n)
    nv = np.column_stack([n, np.sqrt(1 - (n ** 2).sum(1))])
    along = (g * nv).sum(1)[:, None] * nv
    assert np.allclose(out, along + (1 + 2 * eps)[:, None] * (g - along), rtol=1e-13, atol=1e-14)


def test_scalar_tti_with_vertical_axis_against_vti():
    """TTI with axis (0, 0, 1) scales x and y exactly as VTI does, and z by the same clipped s that VTI
    leaves out (VTI keeps pz unscaled, a difference the reference has and this build keeps)."""
    rng = np.random.default_rng(5)
    N = 400
    g = rng.standard_normal((N, 3))
    eps, dl = rng.uniform(0, 0.3, N), rng.uniform(-0.1, 0.15, N)
    v, t = _scalar(VTI, g, eps, dl), _scalar(TTI, g, eps, dl, np.zeros((N, 2)))
    assert np.allclose(t[:, :2], v[:, :2], rtol=1e-13, atol=1e-15)
    s = v[:, 0] / (g[:, 0] * (1 + 2 * eps))
    assert np.allclose(t[:, 2], g[:, 2] * s, rtol=1e-13, atol=1e-15)
    assert np.all(s <= 1 + 1e-15)
    # delta > eps would make s > 1: clipped to 1 (min(s, 1))
    clip = dl > eps
    assert clip.any() and np.allclose(s[clip], 1.0, rtol=0, atol=1e-15)


def test_scalar_tti_rotation_equivariance():
    """TTI about axis n equals TTI about z applied in a frame rotated so that n is z, rotated back."""
    rng = np.random.default_rng(6)
    N = 300
    g = rng.standard_normal((N, 3))
    eps, dl = rng.uniform(0, 0.3, N), rng.uniform(0, 0.15, N)
    n = _unit_axes(N, rng)
    got = _scalar(TTI, g, eps, dl, n)
    nv = np.column_stack([n, np.sqrt(1 - (n ** 2).sum(1))])
    want = np.zeros_like(g)
    rot = np.zeros((N, 3, 3))
    for i in range(N):
        a = np.cross(nv[i], [0.0, 0.0, 1.0])
        s, c = np.linalg.norm(a), nv[i, 2]
        if s < 1e-15:
            rot[i] = np.eye(3)
            continue
        k = a / s
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        rot[i] = np.eye(3) + s * K + (1 - c) * K @ K          # Rodrigues: R n = z
        assert np.allclose(rot[i] @ nv[i], [0, 0, 1], atol=1e-14)
    gr = np.einsum("nij,nj->ni", rot, g)
    out_r = _scalar(TTI, gr, eps, dl, np.zeros((N, 2)))
    want = np.einsum("nji,nj->ni", rot, out_r)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)


def test_scalar_phase_velocity():
    """For a plane wave the operator's symbol k . S(k) / |k|^2 is the squared phase velocity.  TTI gives
    the acoustic qP relation (Alkhalifah 1998, 2000; shear velocity set to zero),
        V^2(th) = 1/2 [E + sqrt(E^2 - 8 (eps - delta) sin^2 th cos^2 th)],  E = (1 + 2 eps) sin^2 th + cos^2 th,
    with th the angle to the symmetry axis; VTI gives 1 + 2 eps horizontally and 1 along the axis."""
    rng = np.random.default_rng(7)
    N = 400
    eps = rng.uniform(0, 0.3, N)
    dl = eps * rng.uniform(-0.5, 1.0, N)       # eps >= delta: no clipping
    n = _unit_axes(N, rng)
    nv = np.column_stack([n, np.sqrt(1 - (n ** 2).sum(1))])
    th = rng.uniform(0, np.pi, N)
    perp = np.cross(nv, rng.standard_normal((N, 3)))
    perp /= np.linalg.norm(perp, axis=1)[:, None]
    kmag = rng.uniform(0.5, 2.0, N)
    k = kmag[:, None] * (np.cos(th)[:, None] * nv + np.sin(th)[:, None] * perp)
    sym = (k * _scalar(TTI, k, eps, dl, n)).sum(1) / kmag ** 2
    s2, c2 = np.sin(th) ** 2, np.cos(th) ** 2
    E = (1 + 2 * eps) * s2 + c2
    assert np.allclose(sym, 0.5 * (E + np.sqrt(E * E - 8 * (eps - dl) * s2 * c2)), rtol=1e-12)
    kh = np.column_stack([rng.standard_normal((N, 2)), np.zeros(N)])
    assert np.allclose((kh * _scalar(VTI, kh, eps, dl)).sum(1) / (kh ** 2).sum(1), 1 + 2 * eps, rtol=1e-13)
    kv = np.column_stack([np.zeros((N, 2)), rng.standard_normal(N)])
    assert np.allclose(_scalar(VTI, kv, eps, dl), kv, rtol=0, atol=0)


# ---- KernelCPU::VTI / TTI: gradient -> scalar -> divergence ---------------------------------------
def _first_derivs(engine, shape, spacing, order):
    ny, nx, nz = shape
    d = {0: spacing[2], 1: spacing[0], 2: spacing[1]}
    if engine == FD:
        Ds = {a: _fd_matrix(shape[ARR_AX[a]], order // 2, d[a]) for a in (0, 1, 2)}
        return Ds, Ds
    grad = {a: _cfd_matrix(shape[ARR_AX[a]], d[a], 0) for a in (0, 1, 2)}
    div = {a: _cfd_matrix(shape[ARR_AX[a]], d[a], 1) for a in (0, 1, 2)}
    return grad, div


@pytest.mark.parametrize("engine,order,model", [(FD, 8, VTI), (FD, 16, TTI), (FD, 8, TTI), (CFD, 10, VTI),
                                                 (CFD, 10, TTI)])
@pytest.mark.parametrize("pml", [False, True])
def test_vti_tti_operator_composition(engine, order, model, pml):
    """KernelCPU::VTI/TTI is pb <- pb - vel * (Dx- Sx + Dz- Sz + Dy- Sy), S = Scalar(Dx+ p, Dy+ p, Dz+ p),
    with each derivative's PML: checked against dense derivative operators and the pinned scalar
    operator, with the first-call PML closed form when the layers are on."""
    shape, spacing = (14, 16, 18), (10.0, 12.0, 9.0)
    nb = [3, 4, 2, 3, 4, 3] if pml else [0] * 6
    dt, vmax, eps_pml, sc = 1e-3, 3000.0, 1e-3, 0.3
    rng = np.random.default_rng(order + model + 10 * pml)
    p0 = smooth_field(shape, -1, 1, 1).astype(np.float64) + 0.1 * rng.standard_normal(shape)
    pb = rng.standard_normal(shape)
    vel = 0.1 + 0.2 * rng.random(shape)
    eps = 0.3 * rng.random(shape)
    dl = 0.15 * rng.random(shape)
    pjx, pjy = (_unit_axes(shape[0] * shape[1] * shape[2], rng).T.reshape((2,) + shape)) if model == TTI else (None, None)
    e = _engine(shape, spacing, engine=engine, order=order)
    e.set_model(vel, eps, dl, None if pjx is None else pjx.copy(), None if pjy is None else pjy.copy())
    e.set_boundary(nb, dt, vmax, eps_pml, True, sc, True, 3)
    got, pr = pb.copy(), np.zeros(shape)
    e.kernel(model, p0, got, pr)
    grad, div = _first_derivs(engine, shape, spacing, order)
    damp = {a: _damping(shape, a, nb, spacing, dt, vmax, eps_pml, sc) if pml else 1.0 for a in (0, 1, 2)}
    gx, gy, gz = (damp[a] * _apply(grad[a], p0, a) for a in (1, 2, 0))
    N = p0.size
    S = _scalar(model, np.column_stack([gx.ravel(), gy.ravel(), gz.ravel()]), eps.ravel(), dl.ravel(),
                None if pjx is None else np.column_stack([pjx.ravel(), pjy.ravel()]))
    sx, sy, sz = (S[:, i].reshape(shape) for i in range(3))
    want = pb - vel * (damp[1] * _apply(div[1], sx, 1) + damp[0] * _apply(div[0], sz, 0) +
                       damp[2] * _apply(div[2], sy, 2))
    assert N == got.size
    assert np.allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_vti_reduces_to_scaled_iso_for_a_field_along_x():
    """Constant eps, delta, no PML, p varying along x only: away from the y/z edges (where the zero halo
    makes Dy p, Dz p nonzero) the VTI operator is (1 + 2 eps) Dx Dx p, the ISO first-derivative operator
    scaled."""
    shape, spacing, M = (20, 24, 22), (10.0, 10.0, 10.0), 4
    e = _engine(shape, spacing)
    vel = np.full(shape, 0.04)
    eps, dl = np.full(shape, 0.2), np.full(shape, 0.07)
    e.set_model(vel, eps, dl)
    e.set_boundary([0] * 6, 1e-3, 3000.0, dim_pml=3)
    x = np.arange(shape[1])
    p0 = np.broadcast_to(np.sin(0.37 * x)[None, :, None], shape).copy()
    got = np.zeros(shape)
    e.kernel(VTI, p0, got, np.zeros(shape))
    Dx = _fd_matrix(shape[1], M, spacing[0])
    want = -vel * 1.4 * _apply(Dx @ Dx, p0, 1)
    c = slice(2 * M, -2 * M)
    assert np.allclose(got[c, :, c], want[c, :, c], rtol=0, atol=1e-13)


# ---- the shot time loop ---------------------------------------------------------------------------
def _shot(cls, nt, spacing=10.0, shape=(20, 22, 24), src=(10.0, 11.0, 12.0), snap_every=1, tmute=0.0):
    ny, nx, nz = shape
    rcv = np.array([[4.0, 5.0, 7.0], [15.0, 6.0, 7.0], [8.5, 16.0, 8.0], [12.0, 12.0, 11.0]], np.float32)
    return cls(nx, ny, nz, spacing, 1e-3, nt, FD, 8, 1, [0] * 6, 2500.0, src, 20.0, 3, 0.3, 0.3, rcv,
               snap_every=snap_every, tmute_frac=tmute, threads=2)


@pytest.mark.parametrize("cls,tol", [(pipeline.Shot64, 1e-11), (pipeline.Shot, 2e-5)])
def test_modelling_and_backpropagation_are_adjoint(cls, tol):
    """Without PML and with a constant velocity the step is symmetric, so the backward loop of migration
    (receiver samples it+1 injected, loop from nt-1 down) is the adjoint of the forward modelling loop
    (source sample it-1 injected, the record extracted after the injection):
        <d, F s> = sum_it s[it] * E_src(v_it),
    with v_it the backward field at step it.  Shifting either side by one step breaks the identity."""
    nt = 60
    shot = _shot(cls, nt)
    shape = shot.shape
    shot.set_model(np.full(shape, 2000.0))
    rng = np.random.default_rng(12)
    s = rng.standard_normal(nt).astype(shot.wavelet.dtype)
    shot.wavelet = s
    Fs = shot.model().astype(np.float64)
    d = rng.standard_normal(Fs.shape)
    ext = pipeline.extract if cls is pipeline.Shot else (lambda w, sh, t: shot._extract(w, t))
    e = np.zeros(nt)

    def visit(it, w0):
        e[it] = float(ext(w0, shape, shot.src)[0])
    shot.backward(d.astype(np.float32) if cls is pipeline.Shot else d, visit)
    lhs = float((d * Fs).sum())
    rhs = float((s.astype(np.float64) * e).sum())
    assert abs(lhs - rhs) <= tol * abs(lhs), (lhs, rhs)
    for shift in (-1, 1):
        wrong = float((s.astype(np.float64) * np.roll(e, shift)).sum())
        assert abs(lhs - wrong) > 1e3 * tol * abs(lhs)


def test_snapshot_cadence_and_time_taper():
    """rtm() correlates the forward field stored after the injection at steps it = k s with the backward
    field at the same step; with a time taper the first ntaperT = rint(tmute nt / k) snapshots are
    weighted by the rising half cosine sin^2(pi s / (2 ntaperT)) (ImagingCondition::calTapes)."""
    nt, k = 48, 3
    shot = _shot(pipeline.Shot64, nt, shape=(20, 22, 18), src=(10.0, 11.0, 8.0), snap_every=k)
    shape = shot.shape
    shot.set_model(np.full(shape, 2000.0) + 300.0 * smooth_field(shape, 0, 1, 3))
    rec = shot.model()
    fwd, bwd = {}, {}
    shot.forward(lambda it, w: fwd.__setitem__(it, w.copy()))
    shot.backward(rec, lambda it, w: bwd.__setitem__(it, w.copy()))
    contrib = {s: fwd[s * k] * bwd[s * k] for s in range(nt // k)}
    img, _ = shot.rtm(rec)
    assert np.allclose(img, sum(contrib.values()), rtol=0, atol=1e-12 * np.abs(img).max())
    tm = 0.25
    shot.tmute_frac = tm
    img_t, _ = shot.rtm(rec)
    ntaper = int(round(tm * nt / k))
    # snapshot 0 is zero (the source sample it - 1 = -1 is not injected), so its weight is not observable
    assert not contrib[0].any()
    A = np.column_stack([contrib[s].ravel() for s in range(1, ntaper)])
    w, *_ = np.linalg.lstsq(A, (img_t - img).ravel(), rcond=None)
    assert np.allclose(w + 1, np.sin(np.pi * np.arange(1, ntaper) / (2 * ntaper)) ** 2, atol=1e-9)


# ---- IRREGULAR z grid (FdEngine_fd::setJacob) ------------------------------------------------------------

def _stretch(nz, a=10.0, b=0.15):
    """Depths z(tau) = a tau + b tau^2 (FP32): the Taylor differences of z are exact in the interior."""
    tau = np.arange(nz, dtype=np.float64)
    return (a * tau + b * tau * tau).astype(np.float32)


@pytest.mark.parametrize("order", [8, 16])
def test_irregular_z_derivatives_exact_on_quadratic_stretch(order):
    """On z(tau) quadratic, f = z^2 is a quartic in tau, which the order-8/16 Taylor stencils differentiate
    exactly: the z Laplacian is 2 and the first derivative of f = z is 1 away from the grid's ends.  A
    wrong sign of the slope term or a missing 1/zp^2 fails by O(1)."""
    from oracle import Engine
    ny, nx, nz = 1, 17, 48
    R = order // 2
    e = Engine(nx, ny, nz, 10.0, 10.0, 10.0, 7, order, np.float64)
    e.set_model(np.ones((ny, nx, nz)))
    e.set_boundary([0] * 6, 1e-3, 3000.0)
    zg = _stretch(nz)
    e.set_zgrid(zg)
    z = zg.astype(np.float64)
    f = np.broadcast_to(z * z, (ny, nx, nz)).copy()
    out = np.zeros_like(f)
    e.laplacian(f, out)
    assert np.abs(out[0, nx // 2, R:nz - R] - 2.0).max() < 1e-6
    g = np.broadcast_to(z, (ny, nx, nz)).copy()
    d = np.zeros_like(g)
    e.deriv1(0, g, d, 0, 1)
    assert np.abs(d[0, nx // 2, R:nz - R] - 1.0).max() < 1e-6
    e.set_zgrid(None)   # the regular 1/dz metric misses the stretch
    e.laplacian(f, out)
    assert np.abs(out[0, nx // 2, R:nz - R] - 2.0).max() > 1.0


def test_uniform_zgrid_reproduces_the_regular_grid():
    """zgrid = dz * iz through the per-iz weights agrees with the regular 1/dz weights to rounding."""
    from oracle import Engine
    ny, nx, nz = 6, 17, 40
    rng = np.random.default_rng(3)
    u = rng.standard_normal((ny, nx, nz))
    outs = []
    for zg in (None, (10.0 * np.arange(nz)).astype(np.float32)):
        e = Engine(nx, ny, nz, 10.0, 10.0, 10.0, 7, 8, np.float64)
        e.set_model(np.ones((ny, nx, nz)))
        e.set_boundary([0] * 6, 1e-3, 3000.0)
        e.set_zgrid(zg)
        lap, d1 = np.zeros_like(u), np.zeros_like(u)
        e.laplacian(u, lap)
        e.deriv1(0, u, d1, 0, 1)
        outs.append((lap, d1))
    for a, b in zip(*outs):
        assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()


def test_compact_jacobz_gives_the_physical_derivative():
    """The compact engine's z derivative on an IRREGULAR grid (scaled by jacobz = dz / zgrid') of a Gaussian
    in physical depth matches its analytic derivative (FP64, 1e-4 of the peak; the regular metric is off by
    O(1)); a uniform zgrid reproduces the regular derivative bit for bit (jacobz = 1)."""
    from oracle import Engine
    ny, nx, nz = 1, 4, 96
    tau = np.arange(nz, dtype=np.float64)
    zg = (10.0 * tau + 0.12 * tau * tau).astype(np.float32)
    z = zg.astype(np.float64)
    zc, s = z[nz // 2], 60.0
    g = np.exp(-((z - zc) / s) ** 2)
    dg = -2.0 * (z - zc) / s ** 2 * g
    e = Engine(nx, ny, nz, 10.0, 10.0, 10.0, 4, 10, np.float64)
    e.set_model(np.ones((ny, nx, nz)))
    e.set_boundary([0] * 6, 1e-3, 3000.0)
    u = np.broadcast_to(g, (ny, nx, nz)).copy()
    out = np.zeros_like(u)
    e.set_zgrid(zg)
    e.deriv1(0, u, out, 0, 1)
    assert np.abs(out[0, 1] - dg).max() < 1e-4 * np.abs(dg).max()
    e.set_zgrid(None)
    e.deriv1(0, u, out, 0, 1)
    assert np.abs(out[0, 1] - dg).max() > 0.5 * np.abs(dg).max()
    f = Engine(nx, ny, nz, 10.0, 10.0, 10.0, 4, 10, np.float32)
    f.set_model(np.ones((ny, nx, nz), np.float32))
    f.set_boundary([0] * 6, 1e-3, 3000.0)
    rng = np.random.default_rng(2)
    v = rng.standard_normal((ny, nx, nz)).astype(np.float32)
    a, b = np.zeros_like(v), np.zeros_like(v)
    f.deriv1(0, v, a, 0, 1)
    f.set_zgrid((10.0 * np.arange(nz)).astype(np.float32))
    f.deriv1(0, v, b, 0, 1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
