"""Sensitivity of the channel objective to the fouling-layer shape.

The fouling height on each wall is a periodic uniform cubic B-spline with
``n_ctrl`` control points over the channel length; the bottom wall moves up
by its height and the top wall down.  ``WallShapeProblem.objective`` builds
the geometry for a parameter vector, initialises the seeded state and runs
the step while accumulating

    J = mean over the last `window` steps of
        pbar(0) - pbar(nX-1) + shear_weight * sum over shear links of tau_BL |link|,

keeping population checkpoints every ``interval`` steps.
``WallShapeProblem.gradient`` runs the adjoint sweep on the GPU and maps
dJ/dy of the wall points to dJ/dtheta.  ``WallObjective`` wraps both in a
``torch.autograd.Function``.

The node types and the boundary-link set are piecewise constant in the wall
shape; the gradient holds them at the built geometry (as a finite difference
does whenever its two evaluations share them).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import channel, lib


def bspline_basis(nX: int, n_ctrl: int, xs: np.ndarray) -> np.ndarray:
    """B[i, j] = N3((xs[i] / dx) - j) for the periodic uniform cubic B-spline
    with knot spacing dx = nX / n_ctrl (N3: the centred cubic B-spline)."""
    dx = nX / n_ctrl
    u = np.asarray(xs, dtype=np.float64)[:, None] / dx - np.arange(n_ctrl)[None, :]
    u = (u + n_ctrl / 2.0) % n_ctrl - n_ctrl / 2.0
    a = np.abs(u)
    return np.where(a < 1.0, (4.0 - 6.0 * a * a + 3.0 * a ** 3) / 6.0,
                    np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))


@dataclass
class WallShapeProblem:
    """A serpentine channel whose walls carry a B-spline fouling layer.

    Defaults are the sensitivity workload: 2048 x 256, height H = 200 with
    walls displaced by 0.1 H sin(2 pi x / 1024) (H + 0.2 H fits the 252 rows
    between the solid edge rows), Re_H = 5000 with nu = 0.005, k-omega on
    (k = 1.5 (0.05 U_b)^2, omega = sqrt(k) / (0.09^(1/4) 0.07 H)),
    64 control points per wall drawn U[0, 4] from the seed, T = 4000 steps,
    J averaged over the last 1000, checkpoints every 64 steps."""
    nX: int = 2048
    nY: int = 256
    H: float = 200.0
    amp: float = 20.0
    period: float = 1024.0
    nu: float = 0.005
    Re: float = 5000.0
    n_ctrl: int = 64
    steps: int = 4000
    window: int = 1000
    interval: int = 64
    shear_weight: float = 0.1
    storage: str = "f32"
    seed: int = 11
    noise: float = 0.01
    komega: bool = True
    frozen_turbulence: bool = False  # adjoint with the eddy viscosity held at its recorded values
    rel_tol: float = 1e-4     # k and omega solves (the reference's defaults)
    abs_tol: float = 1e-7
    max_iters: int = 1000

    def __post_init__(self):
        self.mid = (self.nY - 1) / 2.0
        self.Ub = self.Re * self.nu / self.H
        self.Fval = channel.turbulent_forcing(self.H, self.nu, self.Re)
        self.tau0 = 1.0 / (3.0 * self.nu + 0.5)
        # k = 1.5 (0.05 U_b)^2, omega = sqrt(k) / (0.09^(1/4) 0.07 H)
        self.k_ini = 1.5 * (0.05 * self.Ub) ** 2
        self.omega_ini = math.sqrt(self.k_ini) / (0.09 ** 0.25 * 0.07 * self.H)
        self._lat = None

    def initial_theta(self) -> np.ndarray:
        rng = np.random.Generator(np.random.PCG64(self.seed))
        return rng.uniform(0.0, 4.0, size=2 * self.n_ctrl)

    def _shift(self, x):
        return self.amp * np.sin(2.0 * math.pi * x / self.period)

    def points(self, theta: np.ndarray):
        """Wall points (C0, BL) for theta = [bottom control points, top control points]."""
        theta = np.asarray(theta, dtype=np.float64)
        nc = self.n_ctrl

        def bottom(x):
            return self.mid - self.H / 2.0 + self._shift(x) + bspline_basis(self.nX, nc, x) @ theta[:nc]

        def top(x):
            return self.mid + self.H / 2.0 + self._shift(x) - bspline_basis(self.nX, nc, x) @ theta[nc:]

        return channel.wall_points(self.nX, bottom, top)

    def dy_dtheta(self, C0: np.ndarray) -> np.ndarray:
        """[nN, 2 n_ctrl]: y of each wall point against the control points."""
        nN = C0.shape[0]
        half = nN // 2
        B = bspline_basis(self.nX, self.n_ctrl, C0[:, 0])
        J = np.zeros((nN, 2 * self.n_ctrl))
        J[:half, :self.n_ctrl] = B[:half]
        J[half:, self.n_ctrl:] = -B[half:]
        return J

    def geom_params(self):
        return dict(x_release_pos=10.0, x_stop_pos=float(self.nX - 12), pipe_radius=self.H / 2.0)

    def lattice(self) -> lib.Lattice:
        if self._lat is None:
            self._lat = lib.Lattice(self.nX, self.nY, storage=self.storage, komega=self.komega)
            self._lat.set_params(self.tau0, self.Fval, self.nu, kappa_rel_tol=self.rel_tol,
                                 kappa_abs_tol=self.abs_tol, kappa_max_iters=self.max_iters)
        return self._lat

    def build(self, theta):
        """Builds the geometry for theta and the seeded initial state; returns
        the node types."""
        L = self.lattice()
        C0, BL = self.points(theta)
        L.build_geometry(C0, BL, **self.geom_params())
        nType = L.geometry("nType")
        fluid = (nType.astype(np.int32) & 1) == 0
        rng = np.random.Generator(np.random.PCG64(self.seed + 1))
        D = L.D
        ux = self.Ub * (1.0 + rng.uniform(-self.noise, self.noise, D))
        uy = self.Ub * rng.uniform(-self.noise, self.noise, D)
        L.init_state(np.where(fluid, 1.0, 0.0), np.where(fluid, ux, 0.0), np.where(fluid, uy, 0.0))
        if self.komega:
            L.init_komega(self.k_ini, self.omega_ini)
        self._C0 = C0
        return nType

    def objective(self, theta) -> float:
        self.build(theta)
        self._theta = np.array(theta, dtype=np.float64, copy=True)
        return self.lattice().objective_run(self.steps, self.interval, self.window,
                                            shear_weight=self.shear_weight,
                                            frozen_turbulence=self.frozen_turbulence)

    def gradient(self, theta=None) -> np.ndarray:
        """dJ/dtheta of the last objective() call; with theta given, of the
        objective at theta (the objective run is repeated first when the last
        one was at another theta)."""
        if theta is not None:
            theta = np.asarray(theta, dtype=np.float64)
            last = getattr(self, "_theta", None)
            if last is None or last.shape != theta.shape or not np.array_equal(last, theta):
                self.objective(theta)
        dJdy, _ = self.lattice().adjoint()
        return self.dy_dtheta(self._C0).T @ dJdy

    def close(self):
        if self._lat is not None:
            self._lat.close()
            self._lat = None


def _autograd_function():
    import torch

    class WallObjective(torch.autograd.Function):
        """J(theta) of a WallShapeProblem with the adjoint as its backward."""

        @staticmethod
        def forward(ctx, theta, problem):
            th = theta.detach().to("cpu", torch.float64).numpy().copy()
            J = problem.objective(th)
            ctx.problem = problem
            ctx.theta = th
            return theta.new_tensor(J)

        @staticmethod
        def backward(ctx, gJ):
            # the problem may have evaluated another theta since this forward
            # (a line search): gradient() then repeats this forward's run
            g = torch.as_tensor(ctx.problem.gradient(ctx.theta), dtype=gJ.dtype, device=gJ.device)
            return gJ * g, None

    return WallObjective


def wall_objective(theta, problem: WallShapeProblem):
    """torch entry point: J = wall_objective(theta, problem); J.backward()."""
    return _autograd_function().apply(theta, problem)
