"""Grids, synthetic inputs and run configurations (host side, numpy).

Array conventions follow the reference Fortran exactly: a halo field is
``(0:nx+1, 0:ny+1, 0:nz+1)`` in column-major order (x fastest), held here
as a numpy array of shape ``(nz+2, ny+2, nx+2)`` (C order, so index
``[k, j, i]``); ``cellvalue(1:nx,1:ny,1:nz,1:4)`` is ``(4, nz, ny, nx)``;
1-D metrics are length ``n+2`` arrays indexed ``0..n+1``.

Random inputs use a counter-based generator (splitmix64 of
``seed ^ field_id << 56 ^ linear index``), so any host or the oracle can
reproduce a field without shipping data files.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, field_id: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) for (seed, field_id), element q from counter q."""
    idx = np.arange(n, dtype=np.uint64)
    key = np.uint64(seed) ^ (np.uint64(field_id) << np.uint64(56))
    bits = splitmix64(key ^ idx)
    return (bits >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def gaussian(seed: int, field_id: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller over two counter streams."""
    u1 = uniform01(seed, field_id, n)
    u2 = uniform01(seed, field_id + 64, n)
    u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


# ---------------------------------------------------------------- grids

@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    xl: float
    yl: float
    zl: float
    x: np.ndarray
    x12: np.ndarray
    y: np.ndarray
    y12: np.ndarray
    z: np.ndarray
    z12: np.ndarray
    dx: np.ndarray
    dx12: np.ndarray
    dy: np.ndarray
    dy12: np.ndarray
    dz: np.ndarray
    dz12: np.ndarray

    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    @property
    def npoints(self) -> int:
        return self.nx * self.ny * self.nz


def _uniform_axis(n: int, length: float):
    """Centres and faces of SET_C (nse_dominik.f:563-574)."""
    idx = np.arange(n + 2, dtype=np.float64)
    f12 = length * idx / float(n)
    c = length * (idx - 0.50) / float(n)
    return c, f12


def _spacings(c: np.ndarray, f12: np.ndarray):
    """Metric rules of SET_C (nse_dominik.f:576-604), shared by the uniform
    and the stretched axis."""
    n = len(c) - 2
    d12 = np.zeros(n + 2)
    d = np.zeros(n + 2)
    d12[: n + 1] = c[1: n + 2] - c[: n + 1]
    d12[n + 1] = 2.0 * (f12[n + 1] - c[n + 1])
    d[1: n + 2] = f12[1: n + 2] - f12[: n + 1]
    d[0] = d[1]
    d[n + 1] = d[n]
    return d, d12


def set_c(nx, ny, nz, xl, yl, zl) -> Grid:
    """Uniform grid of SET_C (nse_dominik.f:553-607)."""
    x, x12 = _uniform_axis(nx, xl)
    y, y12 = _uniform_axis(ny, yl)
    z, z12 = _uniform_axis(nz, zl)
    dx, dx12 = _spacings(x, x12)
    dy, dy12 = _spacings(y, y12)
    dz, dz12 = _spacings(z, z12)
    return Grid(nx, ny, nz, xl, yl, zl, x, x12, y, y12, z, z12,
                dx, dx12, dy, dy12, dz, dz12)


def channel_grid(nx, ny, nz, xl=2.0 * math.pi, zl=math.pi, beta=1.8) -> Grid:
    """Channel with walls at y = -1 and y = +1 and tanh clustering.

    Faces y12(j) = tanh(beta (2j/ny - 1)) / tanh(beta) for j = 0..ny,
    centres are face midpoints, the ghost centres and the ghost face
    y12(ny+1) are mirrored about the walls; the spacing rules are those of
    SET_C, so every y stencil sees the same metric arrays the reference
    would.  x and z stay uniform (the spectral Poisson solve needs it).
    """
    x, x12 = _uniform_axis(nx, xl)
    z, z12 = _uniform_axis(nz, zl)
    j = np.arange(ny + 1, dtype=np.float64)
    y12 = np.zeros(ny + 2)
    y12[: ny + 1] = np.tanh(beta * (2.0 * j / ny - 1.0)) / math.tanh(beta)
    y12[0], y12[ny] = -1.0, 1.0
    y12[ny + 1] = 2.0 * y12[ny] - y12[ny - 1]
    y = np.zeros(ny + 2)
    y[1: ny + 1] = 0.5 * (y12[: ny] + y12[1: ny + 1])
    y[0] = 2.0 * y12[0] - y[1]
    y[ny + 1] = 2.0 * y12[ny] - y[ny]
    dx, dx12 = _spacings(x, x12)
    dy, dy12 = _spacings(y, y12)
    dz, dz12 = _spacings(z, z12)
    return Grid(nx, ny, nz, xl, 2.0, zl, x, x12, y, y12, z, z12,
                dx, dx12, dy, dy12, dz, dz12)


# --------------------------------------------------------- configuration

@dataclass
class Config:
    name: str
    nx: int
    ny: int
    nz: int
    re: float
    sc: float
    dt: float
    eta: float
    steps: int
    seed: int
    noise: float
    spheres: list = field(default_factory=list)  # (cx, cy, cz, r)
    # (count, 10) records (cx, cy, cz, a, b, c, qw, qx, qy, qz) of ellipsoids;
    # when set, phi0 is generated on the device (vk_levelset_shapes) and
    # `spheres` is unused
    ellipsoids: np.ndarray | None = None
    xl: float = 2.0 * math.pi
    zl: float = math.pi
    beta: float = 1.8

    def grid(self) -> Grid:
        return channel_grid(self.nx, self.ny, self.nz, self.xl, self.zl, self.beta)


def random_spheres(seed: int, count: int, xl: float, zl: float,
                   rmin: float, rmax: float):
    """Seeded sphere list (cx, cy, cz, r) kept inside the walls."""
    u = uniform01(seed, 40, 4 * count).reshape(count, 4)
    out = []
    for a, b, c, d in u:
        r = rmin + (rmax - rmin) * d
        cy = -1.0 + r + (2.0 - 2.0 * r) * b
        out.append((a * xl, cy, c * zl, r))
    return out


def random_ellipsoids(seed: int, count: int, xl: float, zl: float, amin: float,
                      amax: float) -> np.ndarray:
    """Seeded ellipsoid records (count, 10): centre, semi-axes uniform in
    [amin, amax], a uniformly random orientation (the unit quaternion of
    Shoemake's method from three uniforms), the centre kept inside the
    walls by the largest semi-axis."""
    u = uniform01(seed, 41, 9 * count).reshape(count, 9)
    out = np.zeros((count, 10))
    for q in range(count):
        a, b, c, d, e, f, u1, u2, u3 = u[q]
        ax = amin + (amax - amin) * np.array([d, e, f])
        rmax = float(ax.max())
        cy = -1.0 + rmax + (2.0 - 2.0 * rmax) * b
        s1, s2 = math.sqrt(1.0 - u1), math.sqrt(u1)
        quat = (s1 * math.sin(2.0 * math.pi * u2), s1 * math.cos(2.0 * math.pi * u2),
                s2 * math.sin(2.0 * math.pi * u3), s2 * math.cos(2.0 * math.pi * u3))
        out[q] = (a * xl, cy, c * zl, ax[0], ax[1], ax[2]) + quat
    return out


def make_config(name: str) -> Config:
    """The benchmark configurations (channel64, poisson256, step384,
    adjoint256, large768); eta is 1/K with permeability K = 1e-4."""
    if name == "channel64":
        return Config(name, 64, 65, 64, 2800.0, 0.71, 1e-3, 1e4, 100,
                      0x5EED0001, 0.1, [(math.pi, 0.0, 0.5 * math.pi, 0.25)])
    if name == "poisson256":
        return Config(name, 256, 257, 256, 2800.0, 0.71, 1e-3, 1e4, 1,
                      0x5EED0002, 0.0, [])
    if name == "step384":
        sp = random_spheres(0x5EED0003, 8, 2.0 * math.pi, math.pi, 0.1, 0.3)
        return Config(name, 384, 193, 384, 5600.0, 0.71, 1e-3, 1e4, 200,
                      0x5EED0003, 0.05, sp)
    if name == "adjoint256":
        sp = random_spheres(0x5EED0004, 4, 2.0 * math.pi, math.pi, 0.1, 0.3)
        return Config(name, 256, 129, 256, 2800.0, 0.71, 1e-3, 1e4, 200,
                      0x5EED0004, 0.05, sp)
    if name == "large768":
        # dt 5e-4: at 1e-3 this grid runs at Courant numbers near 0.5, where
        # the explicit AB2 convection of centred differences grows, and the
        # run diverges after about 270 steps in FP64 as in the mixed form
        el = random_ellipsoids(0x5EED0005, 64, 2.0 * math.pi, math.pi, 0.05, 0.2)
        return Config(name, 768, 385, 768, 5600.0, 0.71, 5e-4, 1e4, 400,
                      0x5EED0005, 0.05, [], ellipsoids=el)
    raise ValueError(f"unknown configuration {name!r}")


def small_config(nx, ny, nz, seed=0x5EED00FF, re=2800.0, noise=0.1,
                 spheres=None, eta=1e4, dt=1e-3) -> Config:
    """A channel64-like configuration at a test size."""
    if spheres is None:
        spheres = [(math.pi, 0.0, 0.5 * math.pi, 0.6)]
    return Config(f"small{nx}x{ny}x{nz}", nx, ny, nz, re, 0.71, dt, eta, 10,
                  seed, noise, spheres)


def sphere_levelset(g: Grid, spheres) -> np.ndarray:
    """phi0 = max over spheres of r - |x - c| at cell centres (positive in
    the solid), minimum-image distance in the periodic x and z; the halo is
    zero as readgrid3d leaves it (set_levelset.f:202-214)."""
    phi0 = np.zeros(g.shape)
    if not spheres:
        phi0[1:-1, 1:-1, 1:-1] = -1.0
        return phi0
    X = g.x[1: g.nx + 1][None, None, :]
    Y = g.y[1: g.ny + 1][None, :, None]
    Z = g.z[1: g.nz + 1][:, None, None]
    best = None
    for cx, cy, cz, r in spheres:
        ddx = np.abs(X - cx)
        ddx = np.minimum(ddx, g.xl - ddx)
        ddz = np.abs(Z - cz)
        ddz = np.minimum(ddz, g.zl - ddz)
        val = r - np.sqrt(ddx ** 2 + (Y - cy) ** 2 + ddz ** 2)
        best = val if best is None else np.maximum(best, val)
    phi0[1:-1, 1:-1, 1:-1] = best
    return phi0


FIELDS = ("u", "v", "w", "p", "t",
          "u_v", "v_v", "w_v", "u_d", "v_d", "w_d", "u_a", "v_a", "w_a",
          "u_temp", "v_temp", "w_temp", "u_pre", "v_pre", "w_pre",
          "u_ta", "v_ta", "w_ta", "u_l", "v_l", "w_l",
          "u_vt", "u_dt", "u_t", "t_pre", "t_l",
          "rhocp", "thc", "u_ep2", "v_ep2", "w_ep2", "divu")


# the fields a fused time step needs that initial_state fills with
# non-zero values; the rest of the step's fields start at 0
NONZERO_STEP_FIELDS = ("u", "v", "w", "t", "u_pre", "v_pre", "w_pre", "t_pre", "rhocp", "thc")


def initial_state(cfg: Config, g: Grid | None = None, lean: bool = False) -> dict:
    """All halo fields of program main (nse_dominik.f:27-95) initialised
    for a configuration: laminar u = 1.5 (1 - y^2) plus uniform noise,
    v and w noise (v zero on the wall faces), T linear (1 at y=-1, 0 at
    y=+1), rhoCp = thc = 1, history arrays copied from the fields as the
    reference does before its time loop (nse_dominik.f:246-257).
    cellvalue(...,1) is not filled here: it comes from set_levelset.

    With `lean`, only NONZERO_STEP_FIELDS and phi0 are held (the zero
    fields are left to the device, for grids whose full host state would
    not fit).  phi0 is None for a configuration of ellipsoids: it is
    generated on the device."""
    g = g or cfg.grid()
    nx, ny, nz = g.nx, g.ny, g.nz
    n = g.npoints
    st = {k: np.zeros(g.shape) for k in (NONZERO_STEP_FIELDS if lean else FIELDS)}
    noise = cfg.noise

    def rnd(fid):
        return (2.0 * uniform01(cfg.seed, fid, n) - 1.0).reshape(nz, ny, nx)

    yc = g.y[1: ny + 1][None, :, None]
    st["u"][1:-1, 1:-1, 1:-1] = 1.5 * (1.0 - yc ** 2) + noise * rnd(1)
    st["v"][1:-1, 1:-1, 1:-1] = noise * rnd(2)
    st["v"][:, 0, :] = 0.0
    st["v"][:, ny, :] = 0.0
    st["v"][:, ny + 1, :] = 0.0
    st["w"][1:-1, 1:-1, 1:-1] = noise * rnd(3)
    st["t"][1:-1, 1:-1, 1:-1] = np.broadcast_to(0.5 * (1.0 - yc), (nz, ny, nx))
    st["rhocp"][...] = 1.0
    st["thc"][...] = 1.0
    for a, b in (("u_pre", "u"), ("v_pre", "v"), ("w_pre", "w"), ("t_pre", "t")):
        st[a][...] = st[b]
    st["phi0"] = None if cfg.ellipsoids is not None else sphere_levelset(g, cfg.spheres)
    if not lean:
        st["phi"] = np.zeros(g.shape)
        st["cellvalue"] = np.zeros((4, nz, ny, nx))
    return st
