"""Channel wall descriptions and the named workloads.

``wall_points`` follows the reference's channel recipe (MakeChannel_Test.m:68-113):
bottom-wall points from x = -0.75 in unit steps until x passes nX - 0.5, then
the top wall back from that x to -0.75, with one link between consecutive
points of each wall (the file pair load/lsc0.txt, load/lsbl.txt).  The wall
heights are arbitrary functions of x here, so the same recipe describes flat,
wavy and fouled walls.

Seeded inputs use numpy's PCG64 generator; every seed is part of the workload.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def wall_points(nX: int, bottom, top):
    """Returns (C0 [nN,2] float64, BL [nBL,2] int32) for walls y = bottom(x), y = top(x)."""
    xs = [-0.75]
    while not xs[-1] > nX - 0.5:
        xs.append(xs[-1] + 1.0)
    xb = np.array(xs)
    xt = xb[::-1].copy()
    C0 = np.concatenate([np.stack([xb, np.asarray(bottom(xb), dtype=np.float64)], axis=1),
                         np.stack([xt, np.asarray(top(xt), dtype=np.float64)], axis=1)])
    nb = len(xb)
    BL = [(i, i + 1) for i in range(nb - 1)] + [(i, i + 1) for i in range(nb, 2 * nb - 1)]
    return np.ascontiguousarray(C0), np.asarray(BL, dtype=np.int32)


def write_load_files(directory: str, C0: np.ndarray, BL: np.ndarray):
    """load/lsc0.txt (x y rows, %22.16g like save2file.m) and load/lsbl.txt (0-based pairs)."""
    import os
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "lsc0.txt"), "w") as f:
        for x, y in C0:
            f.write(f"{x:.16g}\t{y:.16g}\n")
    with open(os.path.join(directory, "lsbl.txt"), "w") as f:
        for a, b in BL:
            f.write(f"{a}\t{b}\n")


def gaussian_fouling(nX: int, rng: np.random.Generator, n: int = 16, amp=(0.0, 8.0),
                     width=(20.0, 200.0)):
    """Sum of n periodic Gaussians: height(x) = sum A exp(-d(x,c)^2 / (2 w^2))."""
    A = rng.uniform(*amp, size=n)
    w = rng.uniform(*width, size=n)
    c = rng.uniform(0.0, nX, size=n)

    def h(x):
        x = np.asarray(x, dtype=np.float64)[..., None]
        d = np.abs(x - c)
        d = np.minimum(d, nX - d)
        return np.sum(A * np.exp(-d * d / (2.0 * w * w)), axis=-1)

    return h


@dataclass
class Workload:
    """A lattice, its walls and its physical constants."""
    name: str
    nX: int
    nY: int
    C0: np.ndarray
    BL: np.ndarray
    pipe_radius: float
    tau0: float          # relaxation frequency 1/(3 nu + 0.5)
    nu: float
    Fval: float
    komega: bool = False
    k_ini: float = -1.0
    omega_ini: float = -1.0
    Re: float = 0.0
    noise: float = 0.0   # uniform +-noise on u (absolute) or relative to u_init
    u_init: float = 0.0
    seed: int = 0
    steps: int = 1
    geom: dict = field(default_factory=dict)

    @property
    def XF(self):
        return ((self.nX + 127) // 128) * 128

    @property
    def D(self):
        return self.XF * self.nY

    def geom_params(self):
        p = dict(x_release_pos=10.0, x_stop_pos=float(self.nX - 12), pipe_radius=self.pipe_radius)
        p.update(self.geom)
        return p

    def initial_velocity(self, nType: np.ndarray):
        """Seeded initial macros: rho = 1, u = u_init*(1 + U(-n, n)) or U(-n, n)."""
        rng = np.random.Generator(np.random.PCG64(self.seed + 1))
        fluid = (nType.astype(np.int32) & 0x1) == 0
        rho = np.where(fluid, 1.0, 0.0)
        if self.u_init != 0.0:
            ux = self.u_init * (1.0 + rng.uniform(-self.noise, self.noise, self.D))
            uy = self.u_init * rng.uniform(-self.noise, self.noise, self.D)
        else:
            ux = rng.uniform(-self.noise, self.noise, self.D)
            uy = rng.uniform(-self.noise, self.noise, self.D)
        ux = np.where(fluid, ux, 0.0)
        uy = np.where(fluid, uy, 0.0)
        return rho, ux, uy


def _omega(nu: float) -> float:
    return 1.0 / (3.0 * nu + 0.5)


def flat_channel(nX: int, nY: int, *, tau: float = 0.8, Fval: float = 1e-6, noise: float = 1e-4,
                 seed: int = 1, steps: int = 20000, name: str = "flat") -> Workload:
    """Plane channel with solid rows 0 and nY-1 (walls at y = 0.5 and nY - 1.5)."""
    nu = (tau - 0.5) / 3.0
    C0, BL = wall_points(nX, lambda x: np.full_like(x, 0.5), lambda x: np.full_like(x, nY - 1.5))
    return Workload(name, nX, nY, C0, BL, pipe_radius=(nY - 2) / 2.0, tau0=_omega(nu), nu=nu,
                    Fval=Fval, noise=noise, seed=seed, steps=steps)


def fouled_channel(nX: int, nY: int, *, tau: float = 0.8, Fval: float = 1e-6, noise: float = 1e-4,
                   seed: int = 3, steps: int = 100000, amp=(0.0, 8.0), width=(20.0, 200.0),
                   name: str = "fouled") -> Workload:
    """Plane channel with a seeded Gaussian fouling layer on each wall."""
    nu = (tau - 0.5) / 3.0
    rng = np.random.Generator(np.random.PCG64(seed))
    hb = gaussian_fouling(nX, rng, amp=amp, width=width)
    ht = gaussian_fouling(nX, rng, amp=amp, width=width)
    C0, BL = wall_points(nX, lambda x: 0.5 + hb(x), lambda x: nY - 1.5 - ht(x))
    return Workload(name, nX, nY, C0, BL, pipe_radius=(nY - 2) / 2.0, tau0=_omega(nu), nu=nu,
                    Fval=Fval, noise=noise, seed=seed, steps=steps)


def wavy_channel(nX: int, nY: int, *, H: float, amp: float, period: float, nu: float, Re: float,
                 Fval: float, komega: bool = True, noise: float = 0.01, seed: int = 2,
                 steps: int = 50000, name: str = "wavy") -> Workload:
    """Serpentine channel of height H whose walls are both displaced by amp*sin(2 pi x/period)."""
    mid = (nY - 1) / 2.0

    def shift(x):
        return amp * np.sin(2.0 * math.pi * x / period)

    C0, BL = wall_points(nX, lambda x: mid - H / 2.0 + shift(x), lambda x: mid + H / 2.0 + shift(x))
    Ub = Re * nu / H
    k0 = 1.5 * (0.05 * Ub) ** 2
    w0 = math.sqrt(k0) / (0.09 ** 0.25 * 0.07 * H)
    return Workload(name, nX, nY, C0, BL, pipe_radius=H / 2.0, tau0=_omega(nu), nu=nu, Fval=Fval,
                    komega=komega, k_ini=k0, omega_ini=w0, Re=Re, noise=noise, u_init=Ub, seed=seed,
                    steps=steps)


def turbulent_forcing(H: float, nu: float, Re: float) -> float:
    """Body-force increment for bulk Re_H from Dean's friction law C_f = 0.073 Re_2H^-1/4."""
    Ub = Re * nu / H
    cf = 0.073 * (2.0 * Re) ** -0.25
    return 0.5 * cf * Ub * Ub / (H / 2.0)


def ensemble_members(seed: int = 5, n: int = 48, komega: bool = False) -> list:
    """The batched ensemble: n plane channels, nX from {512, ..., 4096} in
    steps of 512, nY from {64, ..., 512} in steps of 64, Re_H ~ U[2000, 20000],
    each with its own seeded Gaussian fouling layer (channel i uses seed
    seed + i).  The fouling amplitude is scaled by nY / 1024 (U[0, 8 nY/1024])
    so the 16 Gaussians leave the narrowest channels open.  Each channel runs
    laminar BGK at bulk velocity 0.05: nu = 0.05 H / Re_H and the Poiseuille
    force F = 12 nu U / H^2.  With komega, k-omega is on with the turbulent
    forcing for Re_H and k = 1.5 (0.05 U)^2, omega = sqrt(k) / (0.09^(1/4) 0.07 H)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        nX = int(rng.integers(1, 9)) * 512
        nY = int(rng.integers(1, 9)) * 64
        Re = float(rng.uniform(2000.0, 20000.0))
        H = nY - 2.0
        Ub = 0.05
        nu = Ub * H / Re
        r = np.random.Generator(np.random.PCG64(seed + i))
        a = (0.0, 8.0 * nY / 1024.0)
        hb = gaussian_fouling(nX, r, amp=a)
        ht = gaussian_fouling(nX, r, amp=a)
        C0, BL = wall_points(nX, lambda x: 0.5 + hb(x), lambda x: nY - 1.5 - ht(x))
        if komega:
            k0 = 1.5 * (0.05 * Ub) ** 2
            out.append(Workload(f"ens{i}", nX, nY, C0, BL, pipe_radius=H / 2.0, tau0=_omega(nu),
                                nu=nu, Fval=turbulent_forcing(H, nu, Re), komega=True, k_ini=k0,
                                omega_ini=math.sqrt(k0) / (0.09 ** 0.25 * 0.07 * H), Re=Re,
                                noise=0.01, u_init=Ub, seed=seed + i, steps=10000))
            continue
        out.append(Workload(f"ens{i}", nX, nY, C0, BL, pipe_radius=H / 2.0, tau0=_omega(nu),
                            nu=nu, Fval=12.0 * nu * Ub / (H * H), noise=0.01, u_init=Ub,
                            seed=seed + i, steps=10000))
    return out


def config(name: str) -> Workload:
    """The named workloads.

    c1  1024 x 128 plane Poiseuille channel, tau = 0.8, F = 1e-6, u noise +-1e-4, 20,000 steps.
    c2  4096 x 512 k-omega serpentine channel, H = 448, wall displacement 30 sin(2 pi x/1024)
        (0.1 H = 44.8 does not fit 512 rows with solid edge rows: H + 2A must stay <= nY - 4),
        nu = 0.005, Re_H = 10,000, 50,000 steps.
    c3  8192 x 1024 plane channel with a 16-Gaussian fouling layer on each wall, physics of c1,
        100,000 steps (the roofline shape).
    """
    if name == "c1":
        return flat_channel(1024, 128, name="c1")
    if name == "c2":
        nu, Re, H = 0.005, 10000.0, 448.0
        return wavy_channel(4096, 512, H=H, amp=30.0, period=1024.0, nu=nu, Re=Re,
                            Fval=turbulent_forcing(H, nu, Re), name="c2")
    if name == "c3":
        return fouled_channel(8192, 1024, name="c3")
    raise KeyError(name)
