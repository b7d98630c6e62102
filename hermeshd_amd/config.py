"""Run configuration: the runtime equivalent of HERMESHD's compile-time
``input.f90`` / ``params.f90`` switches.

Reference: src/input.f90:1-108, src/params.f90:42-129, tests/*/input_*.f90.
The field and mode layout of the state follows src/params.f90:42-47,163 and
src/initialize.f90:126-194.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field
from typing import Sequence, Tuple

BC_NAMES = {"periodic": 0, "outflow": 1, "wall": 2, "noslip": 3, "inflow": 4}
RIEMANN_NAMES = {"central": 0, "hllc": 1, "roe": 2}
INTEGRATOR_NAMES = {"heun": 2, "shu-osher": 3}
VALID_BASIS = {(4, 2), (8, 2), (10, 3)}

# field indices (0-based) of src/params.f90:42-47
RH, MX, MY, MZ, EN, PXX, PYY, PZZ, PXY, PXZ, PYZ = range(11)
# 13-moment model (nQ 13): traceless stress without pzz (= -pxx - pyy), heat flux
T_PXX, T_PYY, T_PXY, T_PXZ, T_PYZ, T_QX, T_QY, T_QZ = range(5, 13)


def reference_T_floor(TK: float = 300.0, mu: float = 18.0) -> float:
    """T_floor of src/params.f90:78-97: (TK*eV_per_K)/te0 in code units."""
    eV_per_K = 8.61728e-5
    L0, t0, n0 = 1.0, 1.0, 2.5e25
    v0 = L0 / t0
    p0 = mu * 1.67e-27 * n0 * v0 ** 2
    te0 = p0 / n0 / 1.6e-19
    return (TK * eV_per_K) / te0


def reference_tait(mu: float = 39.948):
    """(P_1, P_base) of the Murnaghan-Tait water EOS in code units
    (src/params.f90:97-105): 2.15e9/7.2/p0 and 1.01e5/p0 with
    p0 = mu*1.67e-27*n0*v0^2 (mu of the shipped src/input.f90:69)."""
    p0 = mu * 1.67e-27 * 2.5e25 * 1.0 ** 2
    return 2.15e9 / 7.2 / p0, 1.01e5 / p0


def default_cflm(iquad: int) -> float:
    """set_cflm (src/initialize.f90:104-122)."""
    return {2: 0.14, 3: 0.08, 4: 0.05}[iquad]


@dataclass
class HDConfig:
    nx: int
    ny: int
    nz: int
    nbasis: int = 4
    iquad: int = 2
    nQ: int = 5
    ivis: int = 0
    ieos: int = 1
    riemann: str = "hllc"
    limiter: bool = True
    integrator: str = "heun"
    bc: Tuple[str, str, str, str, str, str] = ("periodic",) * 6
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0
    gamma: float = 1.4
    rh_floor: float = 5.0e-6
    T_floor: float = field(default_factory=reference_T_floor)
    coll: float = 0.0
    colvis: float = 0.0
    cflm: float | None = None
    body_force: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    noise: bool = False
    eta_sd: float = 0.0
    zeta_sd: float = 0.0
    seed: int = 1234
    realization: int = 0
    eos_p1: float = field(default_factory=lambda: reference_tait()[0])
    eos_pbase: float = field(default_factory=lambda: reference_tait()[1])
    # 13-moment model (nQ 13): heat-flux relaxation rate and coll_q * conductivity
    coll_q: float = 0.0
    col_q: float = 0.0

    def __post_init__(self):
        if (self.nbasis, self.iquad) not in VALID_BASIS:
            raise ValueError(f"(nbasis, iquad) = {(self.nbasis, self.iquad)} not in {sorted(VALID_BASIS)}")
        if self.nQ not in (5, 11, 13):
            raise ValueError("nQ must be 5 (hydro fields), 11 (hydro + stress, reference layout) or 13 "
                             "(hydro + traceless stress + heat flux)")
        if self.ivis not in (0, 1, 2):
            raise ValueError("ivis must be 0 (explicit), 1 (semi-implicit stress relaxation) or 2 "
                             "(full 10-moment equations)")
        if self.ivis >= 1 and self.nQ == 5:
            raise ValueError("ivis=1,2 need the stress fields (nQ=11 or 13)")
        if self.nQ == 13 and (self.ivis == 2 or self.noise or self.ieos != 1):
            raise ValueError("the 13-moment model runs with ivis 0 or 1, an ideal gas and no noise")
        if self.ivis == 2 and self.nQ != 11:
            raise ValueError("ivis=2 needs the second moments (nQ=11)")
        if self.ivis == 2 and self.noise:
            raise ValueError("ivis=2 has no stochastic stress")
        if self.noise and self.nQ != 11:
            raise ValueError("stochastic stress needs nQ=11")
        if self.ieos not in (1, 2):
            raise ValueError("ieos must be 1 (ideal gas) or 2 (Murnaghan-Tait water)")
        if self.ieos == 2 and self.riemann == "roe":
            raise ValueError("ieos=2 runs with the HLLC or central flux (the Roe solver is ideal-gas only)")
        if self.riemann not in RIEMANN_NAMES:
            raise ValueError(f"riemann must be one of {list(RIEMANN_NAMES)}")
        if self.integrator not in INTEGRATOR_NAMES:
            raise ValueError(f"integrator must be one of {list(INTEGRATOR_NAMES)}")
        bc = tuple(self.bc)
        if len(bc) != 6 or any(b not in BC_NAMES for b in bc):
            raise ValueError(f"bc must be 6 names from {list(BC_NAMES)}")
        for d in range(3):
            if (bc[2 * d] == "periodic") != (bc[2 * d + 1] == "periodic"):
                raise ValueError("a periodic axis must be periodic on both sides")
        if "inflow" in bc[1:]:
            raise ValueError("inflow is an x-low boundary (x_bc_inflow, src/boundary.f90:571-592)")
        self.bc = bc
        if min(self.nx, self.ny, self.nz) < 1:
            raise ValueError("grid extents must be positive")
        if self.cflm is None:
            self.cflm = default_cflm(self.iquad)

    # ------------------------------------------------------------------
    @property
    def shape(self) -> Tuple[int, int, int, int, int]:
        """Torch/C shape (nbasis, nQ, nz, ny, nx) == Fortran Q(nx,ny,nz,nQ,nbasis)."""
        return (self.nbasis, self.nQ, self.nz, self.ny, self.nx)

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def dxi(self) -> float:
        return self.nx / (self.lx / 2.0 - (-(self.lx / 2.0)))

    @property
    def P_floor(self) -> float:
        return self.T_floor * self.rh_floor

    def replace(self, **kw) -> "HDConfig":
        return dataclasses.replace(self, **kw)

    def as_ints(self) -> dict:
        """Numeric form shared by the C-ABI and the oracle."""
        return dict(
            nx=self.nx, ny=self.ny, nz=self.nz, nbasis=self.nbasis, iquad=self.iquad, nQ=self.nQ,
            ivis=self.ivis, ieos=self.ieos, riemann=RIEMANN_NAMES[self.riemann],
            limiter=int(bool(self.limiter)), integrator=INTEGRATOR_NAMES[self.integrator],
            bc=[BC_NAMES[b] for b in self.bc], noise=int(bool(self.noise)),
            lx=self.lx, ly=self.ly, lz=self.lz, gamma=self.gamma, rh_floor=self.rh_floor,
            T_floor=self.T_floor, coll=self.coll, colvis=self.colvis, cflm=self.cflm,
            body_force=list(self.body_force), eta_sd=self.eta_sd, zeta_sd=self.zeta_sd,
            seed=int(self.seed), realization=int(self.realization),
            eos_p1=float(self.eos_p1), eos_pbase=float(self.eos_pbase),
            coll_q=float(self.coll_q), col_q=float(self.col_q),
        )


# ----------------------------------------------------------------------
# The five workloads, as concrete configurations.
# ----------------------------------------------------------------------
def sod_1d(nx: int = 256) -> HDConfig:
    """1D Sod shock tube: 256 cells on [0,1], linear basis, HLLC, limiter,
    SSP-RK2 at CFL 0.3, outflow ends.  The floors are set well below the
    states (rh_floor 5e-6, P_floor 5e-10) since this workload uses p ~ 1."""
    dx = 1.0 / nx
    return HDConfig(nx=nx, ny=1, nz=1, nbasis=4, iquad=2, nQ=5, riemann="hllc", limiter=True,
                    integrator="heun", bc=("outflow", "outflow", "periodic", "periodic", "periodic", "periodic"),
                    lx=1.0, ly=dx, lz=dx, gamma=1.4, T_floor=1e-4, cflm=0.3)


def vortex_3d(nx: int = 128, ny: int = 128, nz: int = 32) -> HDConfig:
    """Isentropic vortex extruded in z: periodic box of width 10 in x with cubic
    cells ([0,10]^2 x [0,2.5] at 128x128x32), complete-quadratic basis,
    3-point Gauss, HLLC, SSP-RK3.  The reference's dt uses dxi only
    (src/hermeshd.f90:174), so cells must stay cubic for the step to be
    stable in y and z."""
    h = 10.0 / nx
    return HDConfig(nx=nx, ny=ny, nz=nz, nbasis=10, iquad=3, nQ=5, riemann="hllc", limiter=True,
                    integrator="shu-osher", bc=("periodic",) * 6, lx=10.0, ly=h * ny, lz=h * nz,
                    gamma=1.4, T_floor=1e-4)


def poiseuille_2d(nx: int = 64, ny: int = 128, vis: float = 0.01) -> HDConfig:
    """10-moment viscous relaxation channel: rho = T = 1, coll = rho*T/vis,
    colvis = coll*vis (src/initialcon.f90:258-259), body force 1e-3 in x,
    no-slip walls in y, periodic in x, limiter off."""
    coll = 1.0 * 1.0 / vis
    return HDConfig(nx=nx, ny=ny, nz=1, nbasis=4, iquad=2, nQ=11, ivis=1, riemann="hllc", limiter=False,
                    integrator="heun", bc=("periodic", "periodic", "noslip", "noslip", "periodic", "periodic"),
                    lx=float(nx), ly=float(ny), lz=1.0, gamma=5.0 / 3.0, T_floor=1.0, coll=coll,
                    colvis=coll * vis, body_force=(1e-3, 0.0, 0.0))


def fluct_ensemble(vis: float = 0.3, kB: float = 1e-4, sizes=((96, 16), (96, 24), (64, 32)),
                   seed: int = 1234):
    """Fluctuating-hydrodynamics ensemble: periodic boxes of unit cells at
    rho = T = 1, v = 0; 10-moment stress relaxation with coll = rho*T/vis,
    colvis = coll*vis (src/initialcon.f90:258-259) giving viscosity vis;
    stochastic stress with the Landau-Lifshitz amplitude eta_sd =
    sqrt(2 vis kB T) for a fluctuation temperature kB*T (kB << 1 keeps the
    fluctuations linear); linear basis, HLLC, SSP-RK2, limiter off.
    `sizes` lists (count, n) for count boxes of n^3 cells.
    Returns (base configuration, list of box extents)."""
    coll = 1.0 / vis
    base = HDConfig(nx=16, ny=16, nz=16, nbasis=4, iquad=2, nQ=11, ivis=0, riemann="hllc", limiter=False,
                    integrator="heun", bc=("periodic",) * 6, lx=16.0, ly=16.0, lz=16.0, gamma=1.4,
                    T_floor=1e-4, coll=coll, colvis=coll * vis, noise=True,
                    eta_sd=math.sqrt(2.0 * vis * kB), zeta_sd=0.0, seed=seed, realization=0)
    dims = [(n, n, n) for count, n in sizes for _ in range(count)]
    return base, dims


def taylor_green_3d(n: int = 192, Re: float = 1600.0, Pr: float = 0.71, gamma: float = 1.4,
                    riemann: str = "roe") -> HDConfig:
    """13-moment Taylor-Green vortex: n^3 periodic cells on [0, 2 pi]^3 (the
    grid is centred on the origin), complete-quadratic basis, rho = 1,
    mean pressure 100/gamma (Mach 0.1 for |u| = 1), Roe flux, limiter on,
    SSP-RK3.
    Viscosity mu = 1/Re (unit velocity and length): the stress relaxes at
    coll = p/mu with colvis = coll*mu = p; the heat flux relaxes at the same
    rate, coll_q = coll, with conductivity kappa = mu c_p/Pr, c_p =
    gamma/(gamma-1), so col_q = coll_q*kappa.  Semi-implicit relaxation
    (ivis 1): the relaxation times are far below the time step."""
    L = 2.0 * math.pi
    p0 = 100.0 / gamma
    mu = 1.0 / Re
    coll = p0 / mu
    kappa = mu * (gamma / (gamma - 1.0)) / Pr
    return HDConfig(nx=n, ny=n, nz=n, nbasis=10, iquad=3, nQ=13, ivis=1, riemann=riemann, limiter=True,
                    integrator="shu-osher", bc=("periodic",) * 6, lx=L, ly=L, lz=L, gamma=gamma,
                    T_floor=1e-4, coll=coll, colvis=coll * mu, coll_q=coll, col_q=coll * kappa)
