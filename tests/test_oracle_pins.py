"""Pins of the C oracle (CPU): every oracle function is checked against
something the reference's algorithm is fixed by -- closed forms, textbook
solvers written independently (tests/refimpl.py), invariants, or brute
force -- chosen so that a dropped term, wrong sign, swapped index or
transposed table fails.  The reference ships no golden vectors for this
path; its only built-in check is the basis orthonormality test
(src/basis_funcs.f90:634-992), restated in test_basis_orthonormal."""
import math

import numpy as np
import pytest

from hermeshd_amd import ics
from hermeshd_amd.config import EN, MX, MY, MZ, RH, HDConfig, poiseuille_2d, sod_1d, vortex_3d
from tests import refimpl

BASES = [(4, 2), (8, 2), (10, 3)]


def small_cfg(nb, q, **kw):
    base = dict(nx=6, ny=5, nz=4, nbasis=nb, iquad=q, nQ=5, lx=1.2, ly=1.0, lz=0.8, T_floor=1e-4)
    base.update(kw)
    return HDConfig(**base)


# ---------------------------------------------------------------------------
# tables
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nb,q", BASES)
def test_gauss_rule_closed_form(oracle_mod, nb, q):
    o = oracle_mod.Oracle(small_cfg(nb, q).as_ints())
    t = o.tables()
    x, w = refimpl.gauss_legendre(q)
    np.testing.assert_allclose(t["xq"], x, rtol=0, atol=2e-16)
    np.testing.assert_allclose(t["w1"], w, rtol=0, atol=2e-16)
    if q == 2:
        assert t["xq"][1] == 1.0 / math.sqrt(3.0)
    else:
        assert t["xq"][2] == math.sqrt(15.0) / 5.0 and t["w1"][1] == 8.0 / 9.0


@pytest.mark.parametrize("nb,q", BASES)
def test_tables_match_definition(oracle_mod, nb, q):
    """Point orderings and face sign flips against the basis evaluated from its definition."""
    o = oracle_mod.Oracle(small_cfg(nb, q).as_ints())
    t = o.tables()
    xq = t["xq"]
    phi = refimpl.basis_functions(nb)
    for i in range(q):
        for j in range(q):
            for k in range(q):
                g = (i * q + j) * q + k
                want = [f(xq[i], xq[j], xq[k]) for f in phi]
                np.testing.assert_allclose(t["bint"][g], want, rtol=0, atol=1e-15)
    for i in range(q):
        for j in range(q):
            p = i * q + j
            for name, pt in (("bxp", (1, xq[j], xq[i])), ("bxm", (-1, xq[j], xq[i])),
                             ("byp", (xq[i], 1, xq[j])), ("bym", (xq[i], -1, xq[j])),
                             ("bzp", (xq[j], xq[i], 1)), ("bzm", (xq[j], xq[i], -1))):
                np.testing.assert_allclose(t[name][p], [f(*pt) for f in phi], rtol=0, atol=1e-15)
    nf = q * q
    for s, name in enumerate(["bxm", "bxp", "bym", "byp", "bzm", "bzp"]):
        np.testing.assert_array_equal(t["bfaces"][s * nf:(s + 1) * nf], t[name])


@pytest.mark.parametrize("nb,q", BASES)
def test_basis_orthonormal(oracle_mod, nb, q):
    """test_basis_3D (src/basis_funcs.f90:634-992): 0.125*cbasis(r)*sum w3 phi_r phi_s = delta_rs."""
    o = oracle_mod.Oracle(small_cfg(nb, q).as_ints())
    t = o.tables()
    M = 0.125 * t["cbasis"][:, None] * (t["bint"].T * t["w3"]) @ t["bint"]
    np.testing.assert_allclose(M, np.eye(nb), rtol=0, atol=4e-15)
    assert abs(t["w2"].sum() - 4.0) < 1e-15 and abs(t["w3"].sum() - 8.0) < 1e-14


# ---------------------------------------------------------------------------
# point physics
# ---------------------------------------------------------------------------
def random_states(rng, n, nQ=5):
    U = np.zeros((n, nQ))
    rho = rng.uniform(0.2, 2.0, n)
    v = rng.uniform(-2.0, 2.0, (n, 3))
    p = rng.uniform(0.1, 3.0, n)
    U[:, 0] = rho
    U[:, 1:4] = rho[:, None] * v
    U[:, 4] = p / 0.4 + 0.5 * rho * (v ** 2).sum(1)
    return U


def test_euler_flux_closed_form(oracle_mod):
    o = oracle_mod.Oracle(small_cfg(4, 2).as_ints())
    rng = np.random.default_rng(0)
    U = random_states(rng, 50)
    for d in range(3):
        F = o.flux_points(U, d + 1)
        for n in range(50):
            np.testing.assert_allclose(F[n], refimpl.euler_flux(U[n], d, 1.4), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("kind", [1, 2])
def test_riemann_vs_textbook(oracle_mod, kind):
    """HLLC (Batten's formulas) against Toro's HLLC form; Roe (legacy
    Eulderink-Mellema form) against Toro's Roe decomposition; random states,
    each direction.  The legacy Roe averages 0.5*(v^2 + cp*P/rho)
    (dg_3D_hydro_lqc_SEAN.f90, swr(:,:,4)), i.e. half the thermal enthalpy of
    the textbook average; the reference is followed, so Toro's solver is
    evaluated with that enthalpy (h_thermal = 0.5)."""
    cfg = small_cfg(4, 2)
    o = oracle_mod.Oracle(cfg.as_ints())
    rng = np.random.default_rng(1 + kind)
    nf = 4
    for d in range(3):
        for trial in range(40):
            UL = random_states(rng, nf)
            UR = random_states(rng, nf)
            out = o.riemann_face(np.vstack([UL, UR]), d + 1, kind)
            for p in range(nf):
                ref = refimpl.hllc_toro(UL[p], UR[p], d, 1.4) if kind == 1 else refimpl.roe_toro(UL[p], UR[p], d, 1.4, h_thermal=0.5)
                np.testing.assert_allclose(out[p, :5], ref, rtol=1e-11, atol=1e-11)


def test_roe_legacy_enthalpy_differs_from_textbook(oracle_mod):
    """Documents the legacy quirk: the textbook Roe flux (h_thermal = 1) differs."""
    o = oracle_mod.Oracle(small_cfg(4, 2).as_ints())
    rng = np.random.default_rng(5)
    UL, UR = random_states(rng, 4), random_states(rng, 4)
    out = o.riemann_face(np.vstack([UL, UR]), 1, 2)
    ref = refimpl.roe_toro(UL[0], UR[0], 0, 1.4, h_thermal=1.0)
    assert np.abs(out[0, :5] - ref).max() > 1e-3


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_riemann_consistency(oracle_mod, kind):
    """F(U, U) = F(U) for central, HLLC and Roe."""
    o = oracle_mod.Oracle(small_cfg(4, 2).as_ints())
    rng = np.random.default_rng(7)
    U = random_states(rng, 4)
    for d in range(3):
        out = o.riemann_face(np.vstack([U, U]), d + 1, kind)
        F = o.flux_points(U, d + 1)
        np.testing.assert_allclose(out, F, rtol=1e-13, atol=1e-13)


def test_hllc_supersonic_upwind(oracle_mod):
    o = oracle_mod.Oracle(small_cfg(4, 2).as_ints())
    rng = np.random.default_rng(3)
    UL = random_states(rng, 4); UR = random_states(rng, 4)
    for U in (UL, UR):
        U[:, 1] = U[:, 0] * 10.0       # u = +10 >> c
        U[:, 4] += 0.5 * U[:, 0] * 100
    out = o.riemann_face(np.vstack([UL, UR]), 1, 1)
    np.testing.assert_array_equal(out, o.flux_points(UL, 1))


def test_hllc_resolves_contact(oracle_mod):
    """An isolated stationary contact (equal u = 0, p) gives zero mass flux and F = p in momentum."""
    o = oracle_mod.Oracle(small_cfg(4, 2).as_ints())
    UL = np.zeros((4, 5)); UR = np.zeros((4, 5))
    UL[:, 0], UR[:, 0] = 1.0, 0.2
    UL[:, 4] = UR[:, 4] = 1.0 / 0.4
    out = o.riemann_face(np.vstack([UL, UR]), 1, 1)
    np.testing.assert_allclose(out[:, 0], 0.0, atol=1e-15)
    np.testing.assert_allclose(out[:, 1], 1.0, rtol=1e-15)
    np.testing.assert_allclose(out[:, 4], 0.0, atol=1e-15)


# ---------------------------------------------------------------------------
# operator invariants
# ---------------------------------------------------------------------------
def smooth_state(cfg, seed=0, amp=0.1):
    """Random smooth periodic state projected with all modes populated."""
    rng = np.random.default_rng(seed)
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.3, -0.2, 0.1))
    q[1:] = amp * rng.standard_normal(q[1:].shape) * 0.2
    q[0] *= 1.0 + amp * 0.3 * rng.standard_normal(q[0].shape)
    q[0, 4] += 0.5
    return q


@pytest.mark.parametrize("nb,q", BASES)
@pytest.mark.parametrize("riemann", ["central", "hllc", "roe"])
def test_free_stream_preserved(oracle_mod, nb, q, riemann):
    cfg = small_cfg(nb, q, riemann=riemann)
    o = oracle_mod.Oracle(cfg.as_ints())
    u = ics.uniform(cfg, rho=1.3, T=0.7, v=(0.4, -0.3, 0.2))
    r = o.rhs(u)
    assert np.abs(r).max() < 1e-13


def test_walls_preserve_rest(oracle_mod):
    for bc in ("wall", "noslip", "outflow"):
        cfg = small_cfg(4, 2, bc=(bc,) * 6)
        o = oracle_mod.Oracle(cfg.as_ints())
        r = o.rhs(ics.uniform(cfg, rho=1.0, T=1.0))
        assert np.abs(r).max() < 1e-13, bc


@pytest.mark.parametrize("nb,q", BASES)
@pytest.mark.parametrize("riemann", ["central", "hllc", "roe"])
def test_conservation_periodic(oracle_mod, nb, q, riemann):
    """Telescoping faces: the cell-average RHS sums to zero over a periodic grid."""
    cfg = small_cfg(nb, q, riemann=riemann)
    o = oracle_mod.Oracle(cfg.as_ints())
    r = o.rhs(smooth_state(cfg, seed=nb))
    s = r[0].reshape(5, -1).sum(1)
    assert np.abs(s).max() < 1e-11 * np.abs(r[0]).max()


@pytest.mark.parametrize("nb,q", BASES)
def test_rotation_symmetry(oracle_mod, nb, q):
    """Cycling the axes (x->y->z) of the state cycles the RHS: catches any
    direction-specific table, index or flux error."""
    n = 4
    cfg = HDConfig(nx=n, ny=n, nz=n, nbasis=nb, iquad=q, nQ=5, lx=1, ly=1, lz=1, T_floor=1e-4)
    o = oracle_mod.Oracle(cfg.as_ints())
    u = smooth_state(cfg, seed=3)
    # mode permutation under (x,y,z) -> (y,z,x): new(x,y,z) = old(z,x,y)
    if nb == 4:
        perm = [0, 3, 1, 2]
    elif nb == 8:
        perm = [0, 3, 1, 2, 6, 4, 5, 7]  # yz<-xy, zx<-yz, xy<-zx
    else:
        perm = [0, 3, 1, 2, 6, 4, 5, 9, 7, 8]

    def rot(a):
        b = a[perm]                                  # modes
        b = b[:, [0, 3, 1, 2, 4]]                    # momentum components (mx,my,mz) <- (mz,mx,my)
        return np.ascontiguousarray(np.transpose(b, (0, 1, 3, 4, 2)))  # spatial: (z,y,x) <- (y,x,z)
    r1 = o.rhs(u)
    r2 = o.rhs(rot(u))
    np.testing.assert_allclose(r2, rot(r1), rtol=0, atol=1e-12 * np.abs(r1).max())


def test_advection_convergence(oracle_mod):
    """Smooth density wave at constant velocity and pressure: the mode-1 RHS
    approaches the exact cell average of -u.grad(rho) at second order."""
    errs = []
    for n in (8, 16, 32):
        cfg = HDConfig(nx=n, ny=n, nz=2, nbasis=4, iquad=2, nQ=5, lx=1, ly=1, lz=2.0 / n, T_floor=1e-4, limiter=False)
        o = oracle_mod.Oracle(cfg.as_ints())
        xc, yc, zc = ics.cell_centres(cfg)
        h = 1.0 / n
        u, v, p, A = 0.7, -0.4, 1.0, 0.2
        kx, ky = 2 * np.pi, 4 * np.pi
        # L2 projection of rho = 1 + A sin(kx x + ky y) onto {1, x, y} per cell (exact integrals)
        X = xc[None, None, :]; Y = yc[None, :, None]
        sx, sy = np.sin(kx * h / 2) / (kx * h / 2), np.sin(ky * h / 2) / (ky * h / 2)
        avg = 1 + A * sx * sy * np.sin(kx * X + ky * Y)
        # d/dxi projection: 3/2 int_{-1}^{1} xi sin(a + b xi) dxi / 2 ... closed form
        def slope(b):
            return 3 * (np.sin(b) - b * np.cos(b)) / b ** 2
        mx_ = A * slope(kx * h / 2) * sy * np.cos(kx * X + ky * Y)
        my_ = A * sx * slope(ky * h / 2) * np.cos(kx * X + ky * Y)
        q = np.zeros(cfg.shape)
        shp = (cfg.nz, cfg.ny, cfg.nx)
        rho_modes = [avg * np.ones(shp), mx_ * np.ones(shp), my_ * np.ones(shp)]
        for r, rm in enumerate(rho_modes):
            q[r, 0] = rm
            q[r, 1] = rm * u
            q[r, 2] = rm * v
            q[r, 4] = rm * 0.5 * (u * u + v * v)
        q[0, 4] += p / 0.4
        r = o.rhs(q)
        exact = -A * sx * sy * (u * kx + v * ky) * np.cos(kx * X + ky * Y)
        errs.append(np.abs(r[0, 0] - exact).max())
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(orders) > 1.8, (errs, orders)


# ---------------------------------------------------------------------------
# limiter, dt, integrators
# ---------------------------------------------------------------------------
def test_limiter_invariants(oracle_mod):
    cfg = small_cfg(10, 3)
    cfg = cfg.replace(rh_floor=0.05, T_floor=1.0)
    o = oracle_mod.Oracle(cfg.as_ints())
    t = o.tables()
    rng = np.random.default_rng(11)
    q = ics.uniform(cfg, rho=0.5, T=1.0)
    q[1:, :5] = 0.2 * rng.standard_normal(q[1:, :5].shape)
    before = q.copy()
    o.limiter(q)
    np.testing.assert_array_equal(q[0], before[0])                 # averages unchanged
    ratio = np.divide(q[1:, :5], before[1:, :5], out=np.ones_like(q[1:, :5]), where=before[1:, :5] != 0)
    assert ratio.min() >= 0 and ratio.max() <= 1 + 1e-15          # theta in [0,1]
    rho_face = np.einsum("pr,rzyx->pzyx", t["bfaces"][: 6 * 9], q[:, 0])
    assert rho_face.min() >= cfg.rh_floor * (1 - 1e-12)
    # identity on a state that is already positive everywhere
    q2 = smooth_state(cfg, seed=5); q2b = q2.copy()
    o.limiter(q2)
    np.testing.assert_array_equal(q2, q2b)


def test_min_dt_closed_form(oracle_mod):
    cfg = small_cfg(4, 2, cflm=0.2)
    o = oracle_mod.Oracle(cfg.as_ints())
    q = ics.uniform(cfg, rho=2.0, T=0.5, v=(0.3, -0.9, 0.2))
    # cs uses gamma*e (no gamma-1 factor), plus the last cell's cs (src/hermeshd.f90:165,174)
    e = (2.0 * 0.5 / 0.4 + 0.5 * 2.0 * (0.09 + 0.81 + 0.04)) / 2.0 - 0.5 * (0.09 + 0.81 + 0.04)
    cs = math.sqrt(1.4 * e)
    vmag = 0.9 + cs
    want = 0.2 / ((vmag + cs) * cfg.dxi)
    assert abs(o.min_dt(q) - want) < 1e-15 * want


@pytest.mark.parametrize("integrator", ["heun", "shu-osher"])
def test_rk_coefficients_relaxation(oracle_mod, integrator):
    """ivis=1, fluid at rest, uniform stress: each stage divides the stress by
    1+dt*coll, so one step has a closed form that pins the SSP weights."""
    cfg = HDConfig(nx=3, ny=3, nz=3, nbasis=4, iquad=2, nQ=11, ivis=1, coll=7.0, colvis=0.0,
                   integrator=integrator, lx=1, ly=1, lz=1, T_floor=1e-4)
    o = oracle_mod.Oracle(cfg.as_ints())
    q = ics.uniform(cfg)
    q[0, 8] = 0.3  # pxy
    q1, q2 = np.zeros_like(q), np.zeros_like(q)
    dt = 0.01
    o.update(q, q1, q2, dt)
    f = 1 / (1 + dt * 7.0)
    want = 0.3 * (0.5 * (1 + f * f) if integrator == "heun" else 1 / 3 + 2 / 3 * f * (0.75 + 0.25 * f * f))
    np.testing.assert_allclose(q[0, 8], want, rtol=1e-14)


def test_body_force_closed_form(oracle_mod):
    g = 1e-3
    cfg = HDConfig(nx=3, ny=2, nz=2, nbasis=4, iquad=2, nQ=5, body_force=(g, 0, 0), lx=1, ly=1, lz=1, T_floor=1e-4)
    o = oracle_mod.Oracle(cfg.as_ints())
    q = ics.uniform(cfg, rho=1.5, T=1.0)
    r = o.rhs(q)
    np.testing.assert_allclose(r[0, 1], 1.5 * g, rtol=1e-14)
    np.testing.assert_allclose(r[0, 4], 0.0, atol=1e-17)
    assert np.abs(r[0, 0]).max() < 1e-14


def test_sod_against_exact_solution(oracle_mod):
    """L1 density error against the exact Riemann solution at t=0.2 is small
    and falls with resolution."""
    errs = []
    for nx in (128, 256):
        cfg = sod_1d(nx)
        o = oracle_mod.Oracle(cfg.as_ints())
        q = ics.sod(cfg)
        q1, q2 = np.zeros_like(q), np.zeros_like(q)
        t, dt = np.array(0.0), np.array(0.0)
        while t < 0.2:
            o.step(q, q1, q2, t, dt)
        xc, _, _ = ics.cell_centres(cfg)
        rho_ex, ps, us = refimpl.exact_riemann(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4, xc / float(t))
        errs.append(np.abs(q[0, 0, 0, 0] - rho_ex).mean())
    assert errs[1] < 0.012 and errs[1] < 0.75 * errs[0], errs


def test_philox_known_answers(oracle_mod):
    """Known-answer vectors of Philox-4x32-10 (Random123 kat_vectors)."""
    f = oracle_mod.philox4x32_10
    assert f([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert f([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert f([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_noise_normals_are_box_muller_of_the_philox_words(oracle_mod):
    """The single-precision Box-Muller (its own log and sin/cos polynomials)
    against double-precision Box-Muller of the same Philox-4x32-10 words:
    the FP32 uniforms and transforms stay within 5e-6 of it."""
    err = 0.0
    seed, real, step, stream = 1234, 3, 7, 1
    key = [(seed ^ (real * 0x9E3779B9)) & 0xFFFFFFFF, (seed >> 32) ^ real]
    for i in range(4000):
        g = oracle_mod.noise_normals(seed, real, step, stream, i)
        u = oracle_mod.philox4x32_10([i & 0xFFFFFFFF, (i >> 32) | (stream << 24), step, 0], key)
        for h in range(2):
            a, b = (u[2 * h] + 0.5) * 2.0 ** -32, (u[2 * h + 1] + 0.5) * 2.0 ** -32
            r = math.sqrt(-2.0 * math.log(a))
            err = max(err, abs(g[2 * h] - r * math.cos(2 * math.pi * b)), abs(g[2 * h + 1] - r * math.sin(2 * math.pi * b)))
    assert err < 5e-6, err


def test_noise_normals_moments(oracle_mod):
    vals = np.concatenate([oracle_mod.noise_normals(1234, 3, 7, 1, i) for i in range(20000)])
    assert abs(vals.mean()) < 0.02 and abs(vals.std() - 1) < 0.02
    assert abs(np.mean(vals ** 4) - 3) < 0.15


# ---------------------------------------------------------------- noise ----
def _noise_cfg(nb=4, n=(8, 8, 8), **kw):
    q = 3 if nb == 10 else 2
    base = dict(nx=n[0], ny=n[1], nz=n[2], nbasis=nb, iquad=q, nQ=11, ivis=1, coll=10.0, colvis=1.0,
                riemann="hllc", noise=True, eta_sd=0.05, zeta_sd=0.0, lx=float(n[0]), ly=float(n[1]),
                lz=float(n[2]), T_floor=1e-4, limiter=False, seed=99)
    base.update(kw)
    return HDConfig(**base)


def test_noise_stress_covariance(oracle_mod):
    """Landau-Lifshitz covariance <S_ij S_kl> = ae^2 (d_ik d_jl + d_il d_jk - 2/3 d_ij d_kl)
    + 3 az^2 d_ij d_kl, zero mean, traceless without bulk noise."""
    o = oracle_mod.Oracle(_noise_cfg().as_ints())
    n = 40000
    S = o.noise_stress(1, np.arange(n), 1.0, 0.0)
    # deviatoric: traceless up to the single-precision storage of each component
    assert np.all(np.abs(S[:, :3].sum(axis=1)) <= 3 * 2.0 ** -24 * np.abs(S[:, :3]).max(axis=1))
    assert np.array_equal(S, S.astype(np.float32).astype(np.float64))   # stored in FP32
    C = np.cov(S.T)
    want = np.diag([4 / 3, 4 / 3, 4 / 3, 1, 1, 1])
    want[0, 1] = want[1, 0] = want[0, 2] = want[2, 0] = want[1, 2] = want[2, 1] = -2 / 3
    assert np.abs(C - want).max() < 0.04
    assert np.abs(S.mean(axis=0)).max() < 0.02
    Sb = o.noise_stress(4, np.arange(n), 0.0, 1.0)            # bulk part only: az * tr(W)
    assert np.allclose(Sb[:, 0], Sb[:, 1]) and np.allclose(Sb[:, 0], Sb[:, 2])
    assert abs(Sb[:, 0].var() - 3.0) < 0.1 and np.abs(Sb[:, 3:]).max() == 0.0
    # a stream and an id select independent draws
    S2 = o.noise_stress(2, np.arange(n), 1.0, 0.0)
    assert abs(np.corrcoef(S[:, 3], S2[:, 3])[0, 1]) < 0.03


def test_noise_conserves_totals_periodic(oracle_mod):
    """One draw per face point shared by both cells: with periodic boundaries the
    noisy fluxes cancel in the domain totals of every conserved field."""
    for nb in (4, 10):
        cfg = _noise_cfg(nb=nb, n=(5, 4, 3), eta_sd=0.3)
        o = oracle_mod.Oracle(cfg.as_ints())
        q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.2, -0.1, 0.3))
        rng = np.random.default_rng(3)
        q[1:, :5] = 0.02 * rng.standard_normal(q[1:, :5].shape)
        r = o.rhs(q)
        r0 = oracle_mod.Oracle(cfg.replace(noise=False).as_ints()).rhs(q)
        assert np.abs(r - r0).max() > 1e-3                    # the noise is there
        tot = r[0].reshape(cfg.nQ, -1).sum(axis=1)
        assert np.abs(tot[:5]).max() < 1e-11 * np.abs(r[0]).max()


def test_noise_affine_in_amplitude(oracle_mod):
    """The stress enters the fluxes linearly and no solver branch depends on it,
    so rhs(2a) - 2 rhs(a) + rhs(0) = 0."""
    cfg = _noise_cfg(nb=10, n=(4, 3, 3), riemann="hllc")
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.3, 0.1, -0.2))
    rng = np.random.default_rng(8)
    q[1:, :5] = 0.01 * rng.standard_normal(q[1:, :5].shape)
    rs = [oracle_mod.Oracle(cfg.replace(eta_sd=a).as_ints()).rhs(q) for a in (0.0, 0.1, 0.2)]
    assert np.abs(rs[2] - 2 * rs[1] + rs[0]).max() < 1e-12 * np.abs(rs[1] - rs[0]).max() * 10


@pytest.mark.parametrize("nb", [4, 10])
def test_noise_face_average_variance(oracle_mod, nb):
    """Fluid at rest: mode 0 of the momentum rhs is the divergence of the
    face-averaged stress, whose variance must be eta_sd^2/(dV dt) (the
    Landau-Lifshitz amplitude).  Var(rhs_mx) = 2 eta^2/(dV dt) (4/3 dxi^2 + dyi^2 + dzi^2)."""
    n = (16, 16, 16) if nb == 4 else (10, 10, 10)
    cfg = _noise_cfg(nb=nb, n=n, eta_sd=0.2, lx=8.0, ly=16.0, lz=12.0)
    o = oracle_mod.Oracle(cfg.as_ints())
    dt = 0.01
    o.set_noise_dt(dt)
    q = ics.uniform(cfg, rho=1.0, T=1.0)
    dxi, dyi, dzi = n[0] / 8.0, n[1] / 16.0, n[2] / 12.0
    dVi = dxi * dyi * dzi
    want = 2 * cfg.eta_sd ** 2 * dVi / dt * (4 / 3 * dxi ** 2 + dyi ** 2 + dzi ** 2)
    samples = []
    for step in range(6 if nb == 4 else 10):
        o.set_step_index(step)
        samples.append(o.rhs(q)[0, 1].ravel())
    got = np.concatenate(samples).var()
    assert abs(got / want - 1) < 0.08


def test_noise_volume_amplitude(oracle_mod):
    """Mode x of the x-momentum rhs (nb 4, two-point rule) collects the volume
    points and the x, y, z faces with independent draws:
    Var = (3/4)^2 eta^2 dVi/dt [64*4/3 dxi^2 + 128/3 dxi^2 + 32/3 (dyi^2 + dzi^2)]."""
    n = (16, 16, 16)
    cfg = _noise_cfg(nb=4, n=n, eta_sd=0.2, lx=16.0, ly=8.0, lz=32.0)
    o = oracle_mod.Oracle(cfg.as_ints())
    dt = 0.02
    o.set_noise_dt(dt)
    q = ics.uniform(cfg, rho=1.0, T=1.0)
    dxi, dyi, dzi = 1.0, 2.0, 0.5
    want = 0.5625 * cfg.eta_sd ** 2 * dxi * dyi * dzi / dt * (
        64 * 4 / 3 * dxi ** 2 + 128 / 3 * dxi ** 2 + 32 / 3 * (dyi ** 2 + dzi ** 2))
    samples = []
    for step in range(6):
        o.set_step_index(step)
        samples.append(o.rhs(q)[1, 1].ravel())
    got = np.concatenate(samples).var()
    assert abs(got / want - 1) < 0.08


def test_noise_keyed_by_step(oracle_mod):
    """The draw depends on (seed, realization, step): an update advances the step
    index; the same key reproduces the same field."""
    cfg = _noise_cfg(nb=4, n=(4, 4, 4), eta_sd=0.1)
    q = ics.uniform(cfg, rho=1.0, T=1.0)
    o = oracle_mod.Oracle(cfg.as_ints())
    a = o.rhs(q)
    assert np.array_equal(a, o.rhs(q))
    o.update(q.copy(), np.zeros_like(q), np.zeros_like(q), 1e-3)
    b = o.rhs(q)
    assert not np.allclose(a, b)
    o.set_step_index(0)
    assert np.array_equal(a, o.rhs(q))
    o2 = oracle_mod.Oracle(cfg.replace(realization=1).as_ints())
    assert not np.allclose(a, o2.rhs(q))
    o3 = oracle_mod.Oracle(cfg.replace(seed=100).as_ints())
    assert not np.allclose(a, o3.rhs(q))


# ------------------------------------------------- full 10-moment (ivis 2) ----
def _ten_moment_cfg(**kw):
    base = dict(nx=5, ny=4, nz=3, nbasis=4, iquad=2, nQ=11, ivis=2, coll=4.0, colvis=0.0, riemann="hllc",
                lx=1.0, ly=0.8, lz=0.6, T_floor=1e-4, limiter=False)
    base.update(kw)
    return HDConfig(**base)


def _ten_moment_states(rng, n):
    """Random states whose stress fields are second moments E_ij = rho u_i u_j + P_ij
    with a positive-definite P_ij."""
    U = np.zeros((n, 11))
    rho = rng.uniform(0.5, 1.5, n)
    u = rng.uniform(-0.4, 0.4, (n, 3))
    A = rng.standard_normal((n, 3, 3)) * 0.3
    P = np.einsum("nij,nkj->nik", A, A) + np.eye(3) * rng.uniform(0.5, 1.5, (n, 1, 1))
    E = rho[:, None, None] * np.einsum("ni,nj->nij", u, u) + P
    U[:, 0] = rho
    U[:, 1:4] = rho[:, None] * u
    U[:, 4] = np.trace(P, axis1=1, axis2=2) / 2 + 0.5 * rho * (u ** 2).sum(1)
    U[:, 5], U[:, 6], U[:, 7] = E[:, 0, 0], E[:, 1, 1], E[:, 2, 2]
    U[:, 8], U[:, 9], U[:, 10] = E[:, 0, 1], E[:, 0, 2], E[:, 1, 2]
    return U, rho, u, E


def test_ten_moment_flux_matches_tensor_form(oracle_mod):
    """The explicit per-direction 10-moment stress fluxes equal the tensor form
    F_d(E_ij) = u_i E_jd + u_j E_id + u_d E_ij - 2 rho u_i u_j u_d, and the momentum
    flux is E_id with the trace pressure swapped for the gas pressure."""
    cfg = _ten_moment_cfg()
    o = oracle_mod.Oracle(cfg.as_ints())
    rng = np.random.default_rng(21)
    U, rho, u, E = _ten_moment_states(rng, 40)
    idx = {(0, 0): 5, (1, 1): 6, (2, 2): 7, (0, 1): 8, (1, 0): 8, (0, 2): 9, (2, 0): 9, (1, 2): 10, (2, 1): 10}
    for d in range(3):
        F = o.flux_points(U, d + 1)
        np.testing.assert_allclose(F[:, 0], U[:, 1 + d], rtol=1e-14)
        for i in range(3):
            for j in range(i, 3):
                want = (u[:, i] * E[:, j, d] + u[:, j] * E[:, i, d] + u[:, d] * E[:, i, j]
                        - 2 * rho * u[:, i] * u[:, j] * u[:, d])
                np.testing.assert_allclose(F[:, idx[(i, j)]], want, rtol=1e-12, atol=1e-13)
        trP = np.trace(E, axis1=1, axis2=2) - rho * (u ** 2).sum(1)
        p_gas = (cfg.gamma - 1) * (U[:, 4] - 0.5 * rho * (u ** 2).sum(1))
        for i in range(3):
            want = E[:, i, d] + (p_gas - trP / 3 if i == d else 0.0)
            np.testing.assert_allclose(F[:, 1 + i], want, rtol=1e-12, atol=1e-13)


def test_ten_moment_relaxes_to_isotropy(oracle_mod):
    """Uniform state at rest with an anisotropic pressure tensor: the implicit
    update keeps the trace of E and relaxes its deviatoric part by exactly
    1/(1 + dt*coll) per step (no fluxes, zero source at rest)."""
    cfg = _ten_moment_cfg(bc=("periodic",) * 6)
    o = oracle_mod.Oracle(cfg.as_ints())
    q = ics.uniform(cfg, rho=1.0, T=1.0)
    Pd = np.array([1.6, 0.9, 0.5, 0.2, -0.1, 0.05])     # xx, yy, zz, xy, xz, yz
    q[0, 5:] = Pd[:, None, None, None]
    dt, coll = 0.01, cfg.coll
    out = np.zeros_like(q)
    o.euler_step(q, out, dt)
    tr = Pd[:3].sum()
    dev = Pd.copy()
    dev[:3] -= tr / 3
    want = dev / (1 + dt * coll)
    want[:3] += tr / 3
    np.testing.assert_allclose(out[0, 5:, 0, 0, 0], want, rtol=1e-13)
    np.testing.assert_allclose(out[0, :5], q[0, :5], rtol=1e-13, atol=1e-14)


def test_ten_moment_lax_friedrichs_dissipation(oracle_mod):
    """With ivis 2 the face flux of every field carries the local Lax-Friedrichs
    term -0.5*max(|u_d|+c)*(Q_R - Q_L), c from the trace pressure; with the
    central solver nothing overrides it."""
    cfg = _ten_moment_cfg(riemann="central")
    o = oracle_mod.Oracle(cfg.as_ints())
    rng = np.random.default_rng(5)
    UL, rhoL, uL, EL = _ten_moment_states(rng, 4)
    UR, rhoR, uR, ER = _ten_moment_states(rng, 4)
    for d in range(3):
        out = o.riemann_face(np.vstack([UL, UR]), d + 1, 0)
        FL, FR = o.flux_points(UL, d + 1), o.flux_points(UR, d + 1)
        pL = (np.trace(EL, axis1=1, axis2=2) - rhoL * (uL ** 2).sum(1)) / 3
        pR = (np.trace(ER, axis1=1, axis2=2) - rhoR * (uR ** 2).sum(1)) / 3
        c = np.maximum(np.abs(uL[:, d]) + np.sqrt(cfg.gamma * pL / rhoL), np.abs(uR[:, d]) + np.sqrt(cfg.gamma * pR / rhoR))
        np.testing.assert_allclose(out, 0.5 * (FL + FR) - 0.5 * c[:, None] * (UR - UL), rtol=1e-12, atol=1e-13)


def test_ten_moment_conserves_totals(oracle_mod):
    cfg = _ten_moment_cfg(bc=("periodic",) * 6)
    o = oracle_mod.Oracle(cfg.as_ints())
    rng = np.random.default_rng(8)
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.2, -0.1, 0.1))
    U, _, _, _ = _ten_moment_states(rng, cfg.ncell)
    q[0] = U.T.reshape(11, cfg.nz, cfg.ny, cfg.nx)
    q[1:] = 0.01 * rng.standard_normal(q[1:].shape)
    r = o.rhs(q)
    tot = r[0].reshape(11, -1).sum(axis=1)
    assert np.abs(tot[:5]).max() < 1e-12 * np.abs(r[0]).max()


# ------------------------------------------- inflow boundary, solid mask ----
def _pipe_cfg(**kw):
    base = dict(nx=5, ny=4, nz=2, nbasis=4, iquad=2, nQ=11, ivis=1, coll=10.0, colvis=0.5, riemann="hllc",
                bc=("inflow", "outflow", "noslip", "noslip", "periodic", "periodic"), lx=1.0, ly=0.8, lz=0.4,
                T_floor=1e-4, limiter=False)
    base.update(kw)
    return HDConfig(**base)


@pytest.mark.parametrize("riemann", ["hllc", "roe", "central"])
def test_inflow_ghost_equals_a_periodic_neighbour(oracle_mod, riemann):
    """x_bc_inflow: the exterior state of the x-low face is (rho, mx, 0, 0, E)
    from the prescribed state and the stress fields of the interior.  The
    first cell column then sees exactly what it sees in a periodic box whose
    last column holds that state as constant cells (with the first column's
    constant stress): the two right-hand sides agree bitwise there."""
    cfg = _pipe_cfg(riemann=riemann)
    rng = np.random.default_rng(3)
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.2, 0.05, 0.0))
    q[1:, :5] = 0.02 * rng.standard_normal(q[1:, :5].shape)
    q[0, 5:] = 0.01 * rng.standard_normal(q[0, 5:].shape)       # stress: cell constants only
    q[1:, 5:] = 0.0
    inflow = ics.inflow_profile(cfg, 0.3, 1.1, 0.9)
    oa = oracle_mod.Oracle(cfg.as_ints())
    oa.set_inflow(inflow)
    ra = oa.rhs(q)
    cfgb = cfg.replace(bc=("periodic", "periodic", "noslip", "noslip", "periodic", "periodic"))
    qb = q.copy()
    qb[:, :, :, :, -1] = 0.0
    qb[0, RH, :, :, -1] = inflow[RH, 0].astype(float)
    qb[0, MX, :, :, -1] = inflow[MX, 0]
    qb[0, EN, :, :, -1] = inflow[EN, 0]
    qb[0, 5:, :, :, -1] = q[0, 5:, :, :, 0]
    rb = oracle_mod.Oracle(cfgb.as_ints()).rhs(qb)
    np.testing.assert_array_equal(ra[:, :, :, :, 0], rb[:, :, :, :, 0])
    # and the prescribed state matters
    inflow2 = inflow.copy()
    inflow2[RH] *= 1.3
    oa.set_inflow(inflow2)
    assert np.abs(oa.rhs(q)[:, :, :, :, 0] - ra[:, :, :, :, 0]).max() > 1e-6


def test_inflow_requires_its_state(oracle_mod):
    o = oracle_mod.Oracle(_pipe_cfg().as_ints())
    with pytest.raises(oracle_mod.OracleError):
        o.rhs(ics.uniform(_pipe_cfg()))
    with pytest.raises(ValueError):
        _pipe_cfg(bc=("outflow", "inflow", "noslip", "noslip", "periodic", "periodic"))


def test_inflow_profile_is_parabolic():
    cfg = _pipe_cfg(ny=64)
    d = ics.inflow_profile(cfg, 0.3, 1.0, 2.0)
    _, yc, _ = ics.cell_centres(cfg)
    y = yc + cfg.ly / 2
    u = d[MX, 0, 0] / d[RH, 0, 0]
    np.testing.assert_allclose(u, 0.3 * 4 * y * (cfg.ly - y) / cfg.ly ** 2, rtol=1e-14)
    assert abs(u.max() - 0.3) < 1e-3 and u.min() > 0
    np.testing.assert_allclose(d[EN, 0, 0], 2.0 / (cfg.gamma - 1) + 0.5 * u ** 2, rtol=1e-14)
    assert np.all(d[MY] == 0) and np.all(d[MZ] == 0)
    # the same at every face point and z
    assert np.all(d == d[:, :1, :1])


def test_mask_resets_solid_cells_after_each_step(oracle_mod):
    cfg = _pipe_cfg(nx=8, ny=8, nz=1, bc=("periodic",) * 6, limiter=True)
    mask = ics.cylinder_mask(cfg, 0.0, 0.0, 0.2)
    assert 0 < mask.sum() < mask.size
    rng = np.random.default_rng(1)
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.3, 0.0, 0.0))
    q[1:, :5] = 0.01 * rng.standard_normal(q[1:, :5].shape)
    o = oracle_mod.Oracle(cfg.as_ints())
    ref = q.copy()
    q1, q2 = np.zeros_like(q), np.zeros_like(q)
    o.update(ref, q1, q2, 0.01)
    o.set_mask(mask, 1.25)
    got = q.copy()
    o.update(got, q1, q2, 0.01)
    m = np.broadcast_to(mask, q.shape[2:])
    np.testing.assert_array_equal(got[:, :, ~m], ref[:, :, ~m])      # fluid cells: the unmasked update
    assert np.all(got[0, RH][m] == 1.25) and np.all(got[1:, RH][:, m] == 0)
    assert np.all(got[:, MX:MZ + 1][:, :, m] == 0)
    np.testing.assert_array_equal(got[:, EN][:, m], ref[:, EN][:, m])  # energy untouched
    np.testing.assert_array_equal(got[:, 5:][:, :, m], ref[:, 5:][:, :, m])


def test_cylinder_mask_area_and_reference_problems():
    cfg = HDConfig(nx=128, ny=64, nz=1, nbasis=4, iquad=2, nQ=11, ivis=1, lx=4.0, ly=1.0, lz=0.1, T_floor=0.5,
                   bc=("inflow", "outflow", "noslip", "noslip", "periodic", "periodic"))
    q, inflow, mask = ics.pipe_cylinder_2d(cfg, 1)
    dx, dy = cfg.lx / cfg.nx, cfg.ly / cfg.ny
    area = mask.sum() * dx * dy
    assert abs(area - np.pi * (cfg.ly / 8) ** 2) < 0.05 * np.pi * (cfg.ly / 8) ** 2
    xc, yc, _ = ics.cell_centres(cfg)
    jj, ii = np.nonzero(mask[0])
    cx, cy = xc[ii].mean(), yc[jj].mean()
    assert abs(cx - (-cfg.lx / 2 + cfg.ly / 2)) < dx and abs(cy) < dy      # centred ly/2 from the walls
    assert abs((inflow[MX, 0, 0] / inflow[RH, 0, 0]).max() - 1.5) < 1e-2     # version 1: peak 1.5
    assert np.all(q[0, MX:MZ + 1] == 0) and np.allclose(q[0, EN], 0.5 / (cfg.gamma - 1))
    c5, q5, in5, m5 = ics.setup_problem(cfg, 5, vis=0.01)
    assert m5 is None and abs(c5.coll - 0.5 / 0.01) < 1e-12 and abs(c5.colvis - 0.5) < 1e-15
    assert abs((in5[MX, 0, 0] / in5[RH, 0, 0]).max() - 0.3) < 1e-2
    c2, _, _, _ = ics.setup_problem(cfg.replace(ivis=0, nQ=5, coll=3.0), 2, vis=0.01)
    assert c2.coll == 0.0 and c2.colvis == 0.0


def test_hydro_jet_closed_form():
    cfg = HDConfig(nx=8, ny=32, nz=1, nbasis=4, iquad=2, nQ=5, lx=2.0, ly=2.0, lz=0.1, T_floor=0.1)
    q = ics.set_ic(cfg, 0)
    _, yc, _ = ics.cell_centres(cfg)
    ch = np.cosh(20 * yc / (cfg.ly / 2))
    np.testing.assert_allclose(q[0, MX, 0, :, 3], 1.0 / ch, rtol=1e-14)
    assert np.all(np.abs(q[0, MY] * ch[None, :, None] ** 2) <= 0.0025 + 1e-15)
    e = 0.1 / (cfg.gamma - 1) + 0.5 * (q[0, MX] ** 2 + q[0, MY] ** 2)
    np.testing.assert_allclose(q[0, EN], e, rtol=1e-14)
    assert np.all(q[1:] == 0)


# ------------------------------------------------ Murnaghan-Tait water EOS ----
def _water_cfg(**kw):
    base = dict(nx=5, ny=4, nz=3, nbasis=4, iquad=2, nQ=5, ieos=2, riemann="hllc", lx=1.0, ly=0.8, lz=0.6,
                T_floor=1e-4, limiter=False, eos_p1=2.0, eos_pbase=1.0)
    base.update(kw)
    return HDConfig(**base)


def test_water_eos_point_flux(oracle_mod):
    """flux_calc_pnts_r with ieos 2: the pressure is P_1(rho^7.2 - 1) + P_base
    on top of the ideal-gas pressure of the energy (then floored)."""
    cfg = _water_cfg()
    o = oracle_mod.Oracle(cfg.as_ints())
    rng = np.random.default_rng(2)
    U = np.zeros((20, 5))
    U[:, 0] = rng.uniform(0.9, 1.1, 20)
    U[:, 1:4] = rng.uniform(-0.3, 0.3, (20, 3))
    U[:, 4] = rng.uniform(2.0, 3.0, 20)
    for d in range(3):
        F = o.flux_points(U, d + 1)
        rho, m = U[:, 0], U[:, 1:4]
        p = (cfg.gamma - 1) * (U[:, 4] - 0.5 * (m ** 2).sum(1) / rho)
        p = 2.0 * (rho ** 7.2 - 1) + 1.0 + p
        np.testing.assert_allclose(F[:, 1 + d], m[:, d] ** 2 / rho + p, rtol=1e-13)
        np.testing.assert_allclose(F[:, 4], (U[:, 4] + p) * m[:, d] / rho, rtol=1e-13)


def test_water_eos_hllc_consistency_and_stale_density(oracle_mod):
    """HLLC with ieos 2 is consistent (equal states give the physical flux),
    and, as in the reference, every point's wave speeds use 1/rho of the
    face's last right state: changing that density alone changes the flux at
    the other points, with a nonzero contact speed there."""
    cfg = _water_cfg()
    o = oracle_mod.Oracle(cfg.as_ints())
    nf = 4
    rng = np.random.default_rng(7)
    Q = np.zeros((nf, 5))
    Q[:, 0] = rng.uniform(0.95, 1.05, nf)
    Q[:, 1:4] = rng.uniform(-0.2, 0.2, (nf, 3))
    Q[:, 4] = rng.uniform(2.0, 3.0, nf)
    for d in range(3):
        out = o.riemann_face(np.vstack([Q, Q]), d + 1, 1)
        np.testing.assert_allclose(out, o.flux_points(Q, d + 1), rtol=1e-13, atol=1e-14)
    QL = Q.copy()
    QR = Q.copy()
    QR[:, 0] *= 1.02
    QR[:, 1] += 0.05
    base = o.riemann_face(np.vstack([QL, QR]), 1, 1)
    QR2 = QR.copy()
    QR2[-1, 0] *= 1.5
    moved = o.riemann_face(np.vstack([QL, QR2]), 1, 1)
    assert np.abs(moved[0] - base[0]).max() > 1e-8


def test_water_eos_min_dt_closed_form(oracle_mod):
    """get_min_dt with ieos 2: cs = sqrt(7.2 P_1 rho^6.2 + T_floor)
    (src/hermeshd.f90:166), plus the last cell's cs, over dxi."""
    cfg = _water_cfg()
    o = oracle_mod.Oracle(cfg.as_ints())
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.1, -0.2, 0.05))
    rng = np.random.default_rng(3)
    q[0, 0] *= 1 + 0.01 * rng.standard_normal(q[0, 0].shape)
    rho = q[0, 0]
    v = q[0, 1:4] / rho
    cs = np.sqrt(7.2 * 2.0 * rho ** 6.2 + cfg.T_floor)
    vmag = (np.abs(v).max(axis=0) + cs).max()
    want = cfg.cflm / ((vmag + cs.ravel()[-1]) * cfg.dxi)
    assert abs(o.min_dt(q) - want) <= 1e-14 * want


def test_water_eos_conserves_and_keeps_free_stream(oracle_mod):
    cfg = _water_cfg()
    o = oracle_mod.Oracle(cfg.as_ints())
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.2, -0.1, 0.05))
    assert np.abs(o.rhs(q)).max() < 1e-12
    rng = np.random.default_rng(4)
    q[1:] = 0.01 * rng.standard_normal(q[1:].shape)
    q[0, 0] *= 1 + 0.01 * rng.standard_normal(q[0, 0].shape)
    r = o.rhs(q)
    assert np.abs(r[0].reshape(5, -1).sum(axis=1)).max() < 1e-11 * np.abs(r[0]).max()
    with pytest.raises(ValueError):
        _water_cfg(riemann="roe")


# --------------------------------------------------------- field output ----
def _leg(a, x):
    return [np.ones_like(x), x, 1.5 * x ** 2 - 0.5][a]


@pytest.mark.parametrize("nb,iq", [(4, 2), (8, 2), (10, 3)])
def test_sample_evaluates_the_modal_expansion(oracle_mod, nb, iq):
    """output_vtk's point values at the nvtk^3 sub-cell points equal the
    Legendre expansion evaluated there (closed form per mode), and the
    derived fields follow from them."""
    cfg = HDConfig(nx=3, ny=2, nz=2, nbasis=nb, iquad=iq, nQ=11, ivis=1, lx=1.5, ly=1.0, lz=1.0, T_floor=1e-4)
    o = oracle_mod.Oracle(cfg.as_ints())
    rng = np.random.default_rng(nb)
    q = ics.uniform(cfg, rho=1.0, T=1.0, v=(0.2, 0.1, -0.1))
    q[1:] = 0.05 * rng.standard_normal(q[1:].shape)
    q[0, 5:] = rng.uniform(0.5, 1.0, q[0, 5:].shape)
    deg = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]
    deg += [(1, 1, 1)] if nb == 8 else [(2, 0, 0), (0, 2, 0), (0, 0, 2)] if nb == 10 else []
    deg = deg[:nb]
    for nvtk in (1, 2, 3):
        out = o.sample(q, nvtk, (1 << 14) - 1)
        pts = (2 * np.arange(nvtk) + 1 - nvtk) / nvtk
        # reconstruct the expansion on the sub-grid: (f, z', y', x')
        X = np.tile(pts, cfg.nx)[None, None, :]
        Y = np.tile(pts, cfg.ny)[None, :, None]
        Z = np.tile(pts, cfg.nz)[:, None, None]
        rep = lambda a: np.repeat(np.repeat(np.repeat(a, nvtk, 0), nvtk, 1), nvtk, 2)  # noqa: E731
        val = np.zeros((11, cfg.nz * nvtk, cfg.ny * nvtk, cfg.nx * nvtk))
        for r, (a, b, c) in enumerate(deg):
            phi = _leg(a, X) * _leg(b, Y) * _leg(c, Z)
            for f in range(11):
                val[f] += phi * rep(q[r, f])
        rho = val[0]
        u = val[1:4] / rho
        U = val[4] - 0.5 * rho * (u ** 2).sum(0)
        P = (cfg.gamma - 1) * U
        want = [rho, np.log10(rho), u[0], u[1], u[2], P / rho, U, val[4], P * rho ** -(cfg.gamma - 1), P,
                val[5], val[6], val[7]]
        for k, w in enumerate(want):
            np.testing.assert_allclose(out[k], w.astype(np.float32), rtol=2e-6, atol=1e-6)


def test_sample_vorticity_of_a_shear_flow(oracle_mod):
    """rho = 1, m_x = s*eta (y slope s), m_y = t*xi: output_vtk writes
    -2*(dv/dxi - du/deta)*dxi = -2*(t - s)*dxi, i.e. minus the vorticity
    dv/dx - du/dy when dx = dy."""
    cfg = HDConfig(nx=4, ny=4, nz=1, nbasis=4, iquad=2, nQ=5, lx=2.0, ly=2.0, lz=0.5, T_floor=1e-4)
    o = oracle_mod.Oracle(cfg.as_ints())
    q = ics.uniform(cfg, rho=1.0, T=1.0)
    s, t = 0.3, -0.1
    q[2, MX] = s   # y slope of m_x
    q[1, MY] = t   # x slope of m_y
    vort = o.sample(q, 2, 1 << 13)[0]
    np.testing.assert_allclose(vort, -2 * (t - s) * cfg.dxi, rtol=1e-6)
    dudy, dvdx = s * 2 * cfg.ny / cfg.ly, t * 2 * cfg.nx / cfg.lx
    np.testing.assert_allclose(vort, -(dvdx - dudy), rtol=1e-6)


def test_vtr_writer_round_trip(tmp_path):
    from hermeshd_amd.output import read_vtr, write_vtr
    cfg = HDConfig(nx=3, ny=2, nz=2, nbasis=4, iquad=2, nQ=5, lx=3.0, ly=2.0, lz=1.0)
    rng = np.random.default_rng(0)
    rho = rng.random((4, 4, 6)).astype(np.float32)
    vel = rng.random((3, 4, 4, 6)).astype(np.float32)
    path = tmp_path / "out" / "perseus_p0000_t0001.vtr"
    write_vtr(str(path), cfg, {"Density": rho, "Velocity": vel}, nvtk=2)
    back = read_vtr(str(path))
    np.testing.assert_array_equal(back["Density"], rho.ravel())
    np.testing.assert_array_equal(back["Velocity"], np.stack([vel[c].ravel() for c in range(3)], axis=1))
    np.testing.assert_allclose(back["X"], np.linspace(-1.5, 1.5, 7), rtol=1e-6)
    np.testing.assert_allclose(back["Z"], np.linspace(-0.5, 0.5, 5), rtol=1e-6)
    text = path.read_text()
    assert 'WholeExtent="0 6 0 4 0 4"' in text


# ---------------------------------------------------------------------------
# viscous relaxation model, boundary branches, quadratic basis (tests/pins.py;
# test_oracle_mutations.py shows that each fails under a planted mistake)
# ---------------------------------------------------------------------------
def test_stress_rows_converge_to_strain_rate(oracle_mod):
    from tests import pins
    pins.pin_stress_rows(oracle_mod.Oracle)


def test_stress_drives_momentum_and_energy(oracle_mod):
    from tests import pins
    pins.pin_stress_into_momentum(oracle_mod.Oracle)


def test_relaxation_waves_follow_dispersion_relation(oracle_mod):
    from tests import pins
    pins.pin_relaxation_waves(oracle_mod.Oracle)


def test_poiseuille_channel_steady_parabola(oracle_mod):
    from tests import pins
    pins.pin_poiseuille(oracle_mod.Oracle)


def test_boundary_branches_closed_form(oracle_mod):
    from tests import pins
    pins.pin_outflow_branches(oracle_mod.Oracle)


def test_quadratic_basis_third_order(oracle_mod):
    from tests import pins
    pins.pin_nb10_third_order(oracle_mod.Oracle)


# ---------------------------------------------------------------------------
# 13-moment model (no counterpart in the reference; tests/pins.py)
# ---------------------------------------------------------------------------
def test_13m_free_stream_conservation_relaxation(oracle_mod):
    from tests import pins
    pins.pin_13m_invariants(oracle_mod.Oracle)


def test_13m_rotation_symmetry(oracle_mod):
    from tests import pins
    pins.pin_13m_rotation(oracle_mod.Oracle)


def test_13m_waves_follow_dispersion_relation(oracle_mod):
    from tests import pins
    pins.pin_13m_waves(oracle_mod.Oracle)
