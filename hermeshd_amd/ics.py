"""Initial conditions, generated on the host in numpy (FP64).

Like the reference, the generators set mode 1 only (the cell-centre point
value; higher modes are zero), on the grid centred on the origin
(src/initialcon.f90:20-72, src/params.f90:209-227).  Arrays have shape
(nbasis, nQ, nz, ny, nx).
"""
from __future__ import annotations

import math

import numpy as np

from .config import EN, MX, MY, MZ, RH, HDConfig


def cell_centres(cfg: HDConfig):
    """xc, yc, zc of src/params.f90:209-227 for one rank."""
    dx, dy, dz = cfg.lx / cfg.nx, cfg.ly / cfg.ny, cfg.lz / cfg.nz
    # dxi = nx/(lxu-lxd), dx = 1/dxi exactly as init_spatial_params does
    dx = 1.0 / (cfg.nx / (cfg.lx / 2.0 - (-(cfg.lx / 2.0))))
    dy = 1.0 / (cfg.ny / (cfg.ly / 2.0 - (-(cfg.ly / 2.0))))
    dz = 1.0 / (cfg.nz / (cfg.lz / 2.0 - (-(cfg.lz / 2.0))))
    xc = -(cfg.lx / 2.0) + (np.arange(1, cfg.nx + 1) - 0.5) * dx
    yc = -(cfg.ly / 2.0) + (np.arange(1, cfg.ny + 1) - 0.5) * dy
    zc = -(cfg.lz / 2.0) + (np.arange(1, cfg.nz + 1) - 0.5) * dz
    return xc, yc, zc


def background(cfg: HDConfig) -> np.ndarray:
    """Q = 0; rho = rh_floor; E = T_floor*rh_floor/(gamma-1) (src/initialcon.f90:35-37)."""
    q = np.zeros(cfg.shape, dtype=np.float64)
    q[0, RH] = cfg.rh_floor
    q[0, EN] = cfg.T_floor * cfg.rh_floor / (cfg.gamma - 1.0)
    return q


def _set_primitive(q, cfg, dn, vx, vy, vz, pr):
    gm1 = cfg.gamma - 1.0
    q[0, RH] = dn
    q[0, MX] = dn * vx
    q[0, MY] = dn * vy
    q[0, MZ] = dn * vz
    q[0, EN] = pr / gm1 + 0.5 * dn * (vx ** 2 + vy ** 2 + vz ** 2)


def isentropic_vortex(cfg: HDConfig, ver: int = 1, beta: float = 5.0, jitter: float = 0.0,
                      seed: int = 1234) -> np.ndarray:
    """icid 1 (ver 0) / icid 2 (ver 1), src/initialcon.f90:128-183.  With
    jitter > 0 the density is multiplied by 1 + U(-jitter, jitter) (numpy
    PCG64, seeded) after the state is built."""
    gm1, g = cfg.gamma - 1.0, cfg.gamma
    pi = 4.0 * math.atan(1.0)
    vx_amb = 1.0
    vy_amb, Rgsqi = (0.0, 1.0) if ver == 0 else (1.0, 0.5)
    q = background(cfg)
    xc, yc, _ = cell_centres(cfg)
    xp = xc[None, None, :]
    yp = yc[None, :, None]
    r2 = xp ** 2 + yp ** 2
    e1 = np.exp(Rgsqi * (1 - r2))
    dvx = -yp * beta / (2 * pi) * e1
    dvy = xp * beta / (2 * pi) * e1
    dT = -(gm1 * beta ** 2) / (16 * Rgsqi * g * pi ** 2) * np.exp(2 * Rgsqi * (1 - r2))
    shp = (cfg.nz, cfg.ny, cfg.nx)
    vx = np.broadcast_to(vx_amb + dvx, shp)
    vy = np.broadcast_to(vy_amb + dvy, shp)
    vz = np.zeros(shp)
    te = np.broadcast_to(1.0 + dT, shp)
    dn = te ** (1.0 / gm1)
    pr = dn ** g
    _set_primitive(q, cfg, dn, vx, vy, vz, pr)
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        q[0, RH] *= 1.0 + rng.uniform(-jitter, jitter, size=shp)
    return q


def sod(cfg: HDConfig, rho=(1.0, 0.125), p=(1.0, 0.1), ver: int = 0) -> np.ndarray:
    """Sod tube (src/initialcon.f90:190-235) with explicit left/right states:
    x <= 0 (the centre of the box) gets the left state."""
    q = background(cfg)
    xc, yc, _ = cell_centres(cfg)
    shp = (cfg.nz, cfg.ny, cfg.nx)
    s = xc[None, None, :] + (yc[None, :, None] if ver == 1 else 0.0)
    left = np.broadcast_to(s <= 0, shp)
    dn = np.where(left, rho[0], rho[1])
    pr = np.where(left, p[0], p[1])
    z = np.zeros(shp)
    _set_primitive(q, cfg, dn, z, z, z, pr)
    return q


def gaussian_field(shape, corr_cells: float, seed: int, periodic=(True, True, True)) -> np.ndarray:
    """Unit-variance Gaussian random field: white noise (PCG64) convolved with
    a Gaussian kernel of standard deviation corr_cells cells along each axis
    of extent > 1 (FFT, periodic wrap)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal(shape)
    spec = np.fft.fftn(w)
    for ax, n in enumerate(shape):
        if n == 1:
            continue
        k = np.fft.fftfreq(n)
        g = np.exp(-0.5 * (2 * np.pi * k * corr_cells) ** 2)
        sh = [1] * len(shape)
        sh[ax] = n
        spec = spec * g.reshape(sh)
    f = np.real(np.fft.ifftn(spec))
    f -= f.mean()
    return f / f.std()


def channel_flow(cfg: HDConfig, rho: float = 1.0, T: float = 1.0, u_amp: float = 0.1, corr_cells: float = 8.0,
                 seed: int = 1234) -> np.ndarray:
    """Relaxation-model channel: rho, T uniform, velocity u_amp times a
    Gaussian random field with the given correlation length (x and y
    components, independent fields), zero stress."""
    q = background(cfg)
    shp = (cfg.nz, cfg.ny, cfg.nx)
    vx = u_amp * gaussian_field(shp, corr_cells, seed)
    vy = u_amp * gaussian_field(shp, corr_cells, seed + 1)
    vz = np.zeros(shp)
    pr = np.full(shp, rho * T)
    _set_primitive(q, cfg, np.full(shp, rho), vx, vy, vz, pr)
    return q


def uniform(cfg: HDConfig, rho: float = 1.0, T: float = 1.0, v=(0.0, 0.0, 0.0)) -> np.ndarray:
    q = background(cfg)
    shp = (cfg.nz, cfg.ny, cfg.nx)
    one = np.ones(shp)
    _set_primitive(q, cfg, rho * one, v[0] * one, v[1] * one, v[2] * one, rho * T * one)
    return q


def taylor_green(cfg: HDConfig, rho: float = 1.0, noise: float = 1e-4, seed: int = 1234) -> np.ndarray:
    """Taylor-Green vortex on [0, L]^3 (coordinates shifted from the centred
    grid): u = sin x cos y cos z, v = -cos x sin y cos z, w = 0, uniform
    density, and the pressure of the incompressible vortex
    p = 100/gamma + (rho/16)(cos 2x + cos 2y)(cos 2z + 2), which balances the
    initial momentum flux so that no acoustic transient starts (Mach 0.1 for
    |u| = 1); plus noise*N(0,1) on each velocity component."""
    q = background(cfg)
    xc, yc, zc = cell_centres(cfg)
    x = (xc + cfg.lx / 2)[None, None, :]
    y = (yc + cfg.ly / 2)[None, :, None]
    z = (zc + cfg.lz / 2)[:, None, None]
    shp = (cfg.nz, cfg.ny, cfg.nx)
    rng = np.random.Generator(np.random.PCG64(seed))
    u = np.sin(x) * np.cos(y) * np.cos(z) + noise * rng.standard_normal(shp)
    v = -np.cos(x) * np.sin(y) * np.cos(z) + noise * rng.standard_normal(shp)
    w = noise * rng.standard_normal(shp)
    p = 100.0 / cfg.gamma + (rho / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    _set_primitive(q, cfg, np.full(shp, rho), u, v, w, np.broadcast_to(p, shp))
    return q


def hydro_jet(cfg: HDConfig, seed: int = 1234) -> np.ndarray:
    """icid 0, the 2D unstable jet (src/initialcon.f90:79-114): rho = 1,
    p = T_floor, x momentum 1/cosh(20 y/ly_u) and a y momentum perturbation
    0.005*(U - 1/2)/cosh^2(20 y/ly_u).  The reference draws U with Fortran's
    unseeded random_number; here it is numpy PCG64 with the given seed."""
    q = background(cfg)
    _, yc, _ = cell_centres(cfg)
    shp = (cfg.nz, cfg.ny, cfg.nx)
    dn, te = 1.0, cfg.T_floor
    pr = te * dn
    beta, beta2 = 1.0, 0.005
    lyu = cfg.ly / 2.0
    rng = np.random.Generator(np.random.PCG64(seed))
    rnum = rng.uniform(0.0, 1.0, size=shp) - 0.5
    ch = np.cosh(20 * yc / lyu)[None, :, None]
    smx = np.broadcast_to(beta * dn / ch, shp)
    smy = beta2 * rnum / ch ** 2
    q[0, RH] = dn
    q[0, MX] = smx
    q[0, MY] = smy
    q[0, MZ] = 0.0
    q[0, EN] = pr / (cfg.gamma - 1.0) + 0.5 * (smx ** 2 + smy ** 2) / dn
    return q


def inflow_profile(cfg: HDConfig, ux_amb: float, den: float, pres: float) -> np.ndarray:
    """Parabolic x-low inflow state of set_xlobc_inflow
    (src/boundary.f90:595-622): u = 4 ux_amb y'(ly - y')/ly^2 with y' the cell
    centre's height above the bottom wall, the same at every face point.
    Returns (nQ, nface, nz, ny), the layout of the reference's
    Qxlo_ext_def(ny,nz,nface,nQ); fields other than rh, mx, en are 0."""
    _, yc, _ = cell_centres(cfg)
    yp = yc - (-(cfg.ly / 2.0))
    vx = 4 * ux_amb * yp * (cfg.ly - yp) / cfg.ly ** 2
    nf = cfg.iquad ** 2
    d = np.zeros((cfg.nQ, nf, cfg.nz, cfg.ny))
    d[RH] = den
    d[MX] = (den * vx)[None, None, :]
    d[EN] = (pres / (cfg.gamma - 1.0) + 0.5 * den * vx ** 2)[None, None, :]
    return d


def cylinder_mask(cfg: HDConfig, xctr: float, yctr: float, rad: float) -> np.ndarray:
    """Cells whose centre lies in the disc of radius rad about (xctr, yctr),
    in every z plane (cyl_in_2d_pipe_mask, src/boundary.f90:625-648).
    Returns (nz, ny, nx) bool."""
    xc, yc, _ = cell_centres(cfg)
    r2 = (xc[None, :] - xctr) ** 2 + (yc[:, None] - yctr) ** 2
    return np.broadcast_to(r2 <= rad ** 2, (cfg.nz, cfg.ny, cfg.nx)).copy()


def poiseuille_flow(cfg: HDConfig):
    """icid 5 (src/initialcon.f90:244-285): fluid at rest, rho = 1,
    p = T_floor, with the parabolic inflow of peak 0.3.  Returns (q, inflow)."""
    dn, pr = 1.0, cfg.T_floor
    return uniform(cfg, rho=dn, T=pr / dn), inflow_profile(cfg, 3.0e-1, dn, pr)


def pipe_cylinder_2d(cfg: HDConfig, version: int = 0):
    """icid 6 (version 0, peak inflow 0.3) / icid 7 (version 1, 1.5)
    (src/initialcon.f90:289-361): fluid at rest, rho = 1, p = T_floor, the
    parabolic inflow, and a cylinder of radius ly/8 centred ly/2 from the left
    and bottom walls.  Returns (q, inflow, mask)."""
    dn, pr = 1.0, cfg.T_floor
    ux_amb = 0.3 if version == 0 else 1.5
    lxd, lyd = -(cfg.lx / 2.0), -(cfg.ly / 2.0)
    mask = cylinder_mask(cfg, lxd + cfg.ly / 2.0, lyd + cfg.ly / 2.0, cfg.ly / 8.0)
    return uniform(cfg, rho=dn, T=pr / dn), inflow_profile(cfg, ux_amb, dn, pr), mask


def reference_relaxation(cfg: HDConfig, icid: int, vis: float | None = None) -> HDConfig:
    """coll and colvis as set_ic leaves them (src/initialcon.f90:20-72): zero
    for ivis 0, then icid 0, 5, 6 and 7 set coll = rho*T/vis with rho = 1,
    T = T_floor, and colvis = coll*vis (for any ivis)."""
    if icid in (0, 5, 6, 7):
        if vis is None:
            raise ValueError(f"icid {icid} sets coll from the viscosity: give vis")
        coll = 1.0 * cfg.T_floor / vis
        return cfg.replace(coll=coll, colvis=coll * vis)
    if cfg.ivis == 0:
        return cfg.replace(coll=0.0, colvis=0.0)
    return cfg


def set_ic(cfg: HDConfig, icid: int) -> np.ndarray:
    """The reference's icid switch (src/initialcon.f90:41-69): 0 jet, 1/2
    vortex v0/v1, 3/4 Sod v1/v2 (with the reference's p = P_base scaling
    replaced by the explicit states of ``sod``), 5 Poiseuille, 6/7 cylinder
    in a pipe.  Returns the state; ``setup_problem`` also returns the inflow
    state and the cylinder mask of 5-7 and the relaxation parameters."""
    if icid == 0:
        return hydro_jet(cfg)
    if icid == 1:
        return isentropic_vortex(cfg, ver=0)
    if icid == 2:
        return isentropic_vortex(cfg, ver=1)
    if icid == 3:
        return sod(cfg, ver=0)
    if icid == 4:
        return sod(cfg, ver=1)
    if icid == 5:
        return poiseuille_flow(cfg)[0]
    if icid in (6, 7):
        return pipe_cylinder_2d(cfg, icid - 6)[0]
    raise ValueError(f"icid {icid} is not one of the reference's 0-7")


def setup_problem(cfg: HDConfig, icid: int, vis: float | None = None):
    """Everything set_ic sets for icid: (cfg, state, x-low inflow state or
    None, cylinder mask or None).  With vis given, cfg takes the reference's
    coll and colvis (``reference_relaxation``); with vis None it is kept."""
    cfg2 = cfg if vis is None else reference_relaxation(cfg, icid, vis)
    inflow = mask = None
    if icid == 5:
        q, inflow = poiseuille_flow(cfg2)
    elif icid in (6, 7):
        q, inflow, mask = pipe_cylinder_2d(cfg2, icid - 6)
    else:
        q = set_ic(cfg2, icid)
    return cfg2, q, inflow, mask
