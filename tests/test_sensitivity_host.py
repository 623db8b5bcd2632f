"""CPU tests of the wall-shape parameterisation and of the geometry's
sensitivity data (no GPU): the B-spline basis, dy/dtheta, and the
derivative of every boundary-link distance q with respect to the wall-point
heights, checked against central differences of the geometry builder."""
import numpy as np

from komegafoulingmodel_amd import lib
from komegafoulingmodel_amd import sensitivity as S


def test_bspline_basis_is_a_periodic_partition_of_unity():
    x = np.linspace(-3.0, 300.0, 1001)
    B = S.bspline_basis(256, 8, x)
    assert np.allclose(B.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(S.bspline_basis(256, 8, x + 256.0), B, atol=1e-12)
    # the centred cubic B-spline: 2/3 at its knot, 1/6 one knot away
    k = S.bspline_basis(256, 8, np.array([64.0]))[0]
    assert np.allclose(k, [0, 1 / 6, 2 / 3, 1 / 6, 0, 0, 0, 0], atol=1e-15)


def test_wall_points_follow_theta():
    P = S.WallShapeProblem(nX=256, nY=64, H=48, amp=4.8, period=128, n_ctrl=8)
    th = P.initial_theta()
    C0, _ = P.points(th)
    C1, _ = P.points(th + 1e-3 * np.eye(16)[3])
    fd = (C1[:, 1] - C0[:, 1]) / 1e-3
    assert np.allclose(fd, P.dy_dtheta(C0)[:, 3], atol=1e-9)


def test_link_distance_derivatives_match_central_differences():
    P = S.WallShapeProblem(nX=256, nY=64, H=48, amp=4.8, period=128, n_ctrl=8)
    C0, BL = P.points(0.5 * P.initial_theta())
    L = lib.Lattice(P.nX, P.nY, host_only=True)
    L.build_geometry(C0, BL, **P.geom_params())
    ibb0 = L.geometry("ibb")
    seg = L.geometry("ibbSeg")
    dq = L.geometry("ibbDqDy").reshape(-1, 2)
    assert len(seg) == len(ibb0) and np.all(seg >= 0)
    h = 1e-7
    checked = 0
    for pi in np.random.default_rng(0).choice(C0.shape[0], 40, replace=False):
        qs = []
        for sgn in (1.0, -1.0):
            C = C0.copy()
            C[pi, 1] += sgn * h
            L.build_geometry(C, BL, **P.geom_params())
            if not np.array_equal(L.geometry("ibb"), ibb0):
                break
            qs.append(L.geometry("ibbQ"))
        if len(qs) < 2:
            continue
        fd = (qs[0] - qs[1]) / (2 * h)
        an = np.where(seg == pi, dq[:, 0], 0.0) + np.where(seg + 1 == pi, dq[:, 1], 0.0)
        assert np.max(np.abs(fd - an)) < 1e-6, pi
        checked += 1
    assert checked >= 30


def test_wall_field_gradient_matches_central_differences():
    """kofm_geometry_field_gradient (the wall-distance and wall-spacing part
    of the k-omega adjoint) against central differences of the geometry
    build: for random weights wb, db on the fluid nodes,
    d/dy_p [sum wb WallD + sum db dX] from rebuilt geometries."""
    P = S.WallShapeProblem(nX=256, nY=64, H=48, amp=4.8, period=128, n_ctrl=8)
    C0, BL = P.points(0.5 * P.initial_theta())
    L = lib.Lattice(P.nX, P.nY, host_only=True)
    L.build_geometry(C0, BL, **P.geom_params())
    nT = L.geometry("nType")
    fluid = (nT.astype(np.int32) & 1) == 0
    rng = np.random.default_rng(3)
    wb = np.where(fluid, rng.uniform(-1, 1, fluid.size), 0.0)
    db = np.where(np.tile(fluid, 4), rng.uniform(-1, 1, 4 * fluid.size), 0.0)
    g = L.geometry_field_gradient(wb, db)
    # the distance formula cancels terms of order x * y ~ 1e4 (round-off
    # ~1e-12 in WallD), so the step cannot be much smaller
    h = 1e-5
    checked = nonzero = 0

    def diff(pi, step):
        vals = []
        for sgn in (1.0, -1.0):
            C = C0.copy()
            C[pi, 1] += sgn * step
            L.build_geometry(C, BL, **P.geom_params())
            if not np.array_equal(L.geometry("nType"), nT):
                return None
            vals.append((L.geometry("wallD"), L.geometry("dX")))
        # difference the fields first: only nodes near the moved point change
        return (np.dot(wb, vals[0][0] - vals[1][0]) + np.dot(db, vals[0][1] - vals[1][1])) / (2 * step)

    for pi in rng.choice(C0.shape[0], 60, replace=False):
        fd, fd4 = diff(pi, h), diff(pi, h / 4)
        if fd is None or fd4 is None:
            continue
        # WallD is a minimum over links: where the nearest link changes
        # within +-h the difference is not a derivative
        if abs(fd - fd4) > 1e-4 * max(1.0, abs(fd)):
            continue
        assert abs(fd - g[pi]) < 1e-5 * max(1.0, abs(fd)), (pi, fd, g[pi])
        checked += 1
        nonzero += abs(fd) > 1e-3
    assert checked >= 40 and nonzero >= 20
