"""Pins of the linear-response model of tests/fdt_linear.py (CPU, oracle).

The model predicts the structure factors that the GPU ensemble test checks.
These tests show that its pieces are those of the oracle's scheme:
  - the one-step stencil R reproduces an oracle step of a random
    perturbation;
  - the noise stencils G with the stress covariances C reproduce an oracle
    step driven by the real Philox noise field, at an amplitude small enough
    for the step to be linear;
  - the closed-form transient equals the covariance recursion iterated step
    by step;
  - the wavevector symmetries used to shrink the work hold.
It also records which Riemann solvers make the scheme smooth at rest, the
condition under which the linearisation is exact.
"""
import math

import numpy as np
import pytest

from hermeshd_amd import ics
from hermeshd_amd.config import fluct_ensemble
from tests import fdt_linear as fl


def _cfg(n, riemann="roe", kB=1e-4):
    base, _ = fluct_ensemble(sizes=((1, n),), kB=kB)
    return base.replace(nx=n, ny=n, nz=n, lx=float(n), ly=float(n), lz=float(n), riemann=riemann)


@pytest.fixture(scope="module")
def roe_model(oracle_mod):
    oracle_mod.set_threads(1)
    return fl.LinearModel(_cfg(8), 0.1)


def _odd_defect(oracle_mod, riemann):
    cfg = _cfg(6, riemann)
    o = oracle_mod.Oracle(cfg.as_ints())
    q0 = ics.uniform(cfg)
    q1, q2 = o.zeros(), o.zeros()
    o.set_noise_impulse(1, 0, 0, 0.0)
    d = 1e-7 * np.random.default_rng(3).standard_normal(q0.shape)
    a, b = q0 + d, q0 - d
    o.update(a, q1, q2, 0.1)
    o.update(b, q1, q2, 0.1)
    return np.abs((a - q0) + (b - q0)).max() / np.abs(a - q0).max()


def test_smooth_at_rest_except_hllc(oracle_mod):
    """One step from rest is odd in the perturbation (linear) for Roe and the
    central flux, but not for HLLC: its star fluxes F_K + s_K (Q*_K - Q_K)
    carry the stress of side K while the star pressure leaves it out, so the
    momentum flux jumps by the stress difference when the contact speed
    changes sign."""
    assert _odd_defect(oracle_mod, "roe") < 1e-5
    assert _odd_defect(oracle_mod, "central") < 1e-5
    assert _odd_defect(oracle_mod, "hllc") > 0.05


def _conv(stencil, field):
    """sum_o stencil[o] field(r - o) on a periodic box; stencil (5,5,5,a,b), field (b,z,y,x)."""
    out = np.zeros((stencil.shape[3],) + field.shape[1:])
    for oz in range(5):
        for oy in range(5):
            for ox in range(5):
                sh = np.roll(field, (oz - 2, oy - 2, ox - 2), axis=(1, 2, 3))
                out += np.einsum("ab,bzyx->azyx", stencil[oz, oy, ox], sh)
    return out


def test_stencil_reproduces_oracle_step(oracle_mod, roe_model):
    m, n = roe_model, 8
    cfg = _cfg(n)
    o = oracle_mod.Oracle(cfg.as_ints())
    q0 = ics.uniform(cfg)
    q1, q2 = o.zeros(), o.zeros()
    o.set_noise_impulse(1, 0, 0, 0.0)
    d = 1e-7 * np.random.default_rng(5).standard_normal(q0.shape)
    q = q0 + d
    o.update(q, q1, q2, 0.1)
    nd = m.nb * m.nQ
    act = (q - q0).reshape(nd, n, n, n)
    pred = _conv(m.R, d.reshape(nd, n, n, n))
    assert np.abs(act - pred).max() < 1e-5 * np.abs(act).max()


def test_noise_stencil_reproduces_oracle_step(oracle_mod, roe_model):
    """The drawn noise of one step, pushed through G, gives the oracle's
    step from rest.  kB = 1e-12 keeps the step linear.  The point amplitude
    eta_sd sqrt(dVi/dt) kappa is computed here from the quadrature weights, so
    a wrong kappa or a wrong face/volume id would show."""
    m, n, dt = roe_model, 8, 0.1
    cfg = _cfg(n, kB=1e-12)
    o = oracle_mod.Oracle(cfg.as_ints())
    q0 = ics.uniform(cfg)
    q1, q2 = o.zeros(), o.zeros()
    o.set_step_index(11)
    q = q0.copy()
    o.update(q, q1, q2, dt)
    nd = m.nb * m.nQ
    act = (q - q0).reshape(nd, n, n, n)
    t = o.tables()
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pred = np.zeros_like(act)
    o.set_step_index(11)
    for it, (stream, p, _) in enumerate(m.taus):
        w = t["w3"] if stream == 4 else t["w2"]
        amp = math.sqrt(1.0 / dt) / math.sqrt(np.sum((w / w.sum()) ** 2))
        if stream == 4:
            ids = p * n ** 3 + (ii + n * (jj + n * kk))
        else:
            ids = p * (n + 1) ** 3 + (ii + (n + 1) * (jj + (n + 1) * kk))
        S = o.noise_stress(stream, ids.ravel(), cfg.eta_sd * amp, cfg.zeta_sd * amp).reshape(n, n, n, 6)
        pred += _conv(m.G[it], np.moveaxis(S, -1, 0))
    assert np.abs(act).max() > 0
    assert np.abs(act - pred).max() < 1e-4 * np.abs(act).max()


def test_stress_covariance_of_model(roe_model):
    """C of a face point: Var(S_xy) = ae^2, Var(S_xx) = 4/3 ae^2,
    Cov(S_xx, S_yy) = -2/3 ae^2 (Landau-Lifshitz, no bulk noise), with
    ae^2 = eta_sd^2 (dVi/dt) kappa^2 and kappa^2 = 4 for the 2x2 face rule."""
    m = roe_model
    ae2 = m.cfg.eta_sd ** 2 * (1.0 / 0.1) * 4.0
    C = m.C[0]
    assert C[3, 3] == pytest.approx(ae2, rel=1e-12)
    assert C[0, 0] == pytest.approx(4.0 / 3.0 * ae2, rel=1e-12)
    assert C[0, 1] == pytest.approx(-2.0 / 3.0 * ae2, rel=1e-12)
    assert C[3, 4] == 0.0


def test_transient_equals_recursion(roe_model):
    m = roe_model
    k = np.array([[0.4, 0.0, 0.0], [0.2, 0.3, 0.1]])
    A, Q = m.symbols(k)
    P = np.zeros_like(A)
    acc = np.zeros_like(A)
    equil, nsamp, every = 40, 6, 5
    for s in range(1, equil + nsamp * every + 1):
        P = A @ P @ A.conj().transpose(0, 2, 1) + Q
        if s > equil and (s - equil) % every == 0:
            acc += P
    acc /= nsamp
    Srr, SuL, SuT = m.structure_factors(k, 1e-4, equil, nsamp, every)
    scale = m.dV / 1e-4
    assert Srr == pytest.approx(acc[:, 0, 0].real * scale, rel=1e-10)
    Pm = acc[:, 1:4, 1:4].real * scale
    kn = k / np.linalg.norm(k, axis=1)[:, None]
    uL = np.einsum("ki,kij,kj->k", kn, Pm, kn)
    assert SuL == pytest.approx(uL, rel=1e-10)
    assert SuT == pytest.approx((np.trace(Pm, axis1=1, axis2=2) - uL) / 2, rel=1e-10)


def test_cubic_symmetry(roe_model):
    """Permuting and reflecting the components of k leaves the predicted
    structure factors unchanged (binned_prediction evaluates one wavevector
    per orbit)."""
    m = roe_model
    k = np.array([0.5, 0.3, 0.1])
    ks = np.array([k, k[[1, 2, 0]], k[[2, 0, 1]] * [-1, 1, 1], k[[0, 2, 1]] * [1, -1, -1]])
    S = np.array(m.structure_factors(ks, 1e-4, 30, 3, 5))
    assert np.allclose(S, S[:, :1], rtol=1e-9)


def test_low_k_limits(roe_model):
    """Stationary predictions for a smooth flux as k -> 0.  The sound channel
    reaches the fluctuation-dissipation value S_uL = 1.  The shear channel
    stays at 0.60: the face noise, whose variance gives the face-averaged
    stress the Landau-Lifshitz value, forces the slow shear mode only a third
    as much as a fluctuation-dissipation balance needs (tripling the face
    variance gives S_uT = 1 at k -> 0)."""
    m = roe_model
    _, SuL, SuT = m.structure_factors(np.array([[0.02, 0.0, 0.0]]), 1e-4, 0, 0, 0)
    assert SuL[0] == pytest.approx(1.0, abs=2e-3)
    assert SuT[0] == pytest.approx(0.60, abs=5e-3)


def test_wavevector_counts():
    """box_wavevectors counts the nonzero FFT wavevectors of a box below kmax."""
    wv = fl.box_wavevectors([(8, 8, 8)], kmax=10.0)[8]
    assert sum(wv.values()) == 8 ** 3 - 1
    wv = fl.box_wavevectors([(16, 16, 16)], kmax=0.6)[16]
    assert wv == {(1, 0, 0): 6, (1, 1, 0): 12}
