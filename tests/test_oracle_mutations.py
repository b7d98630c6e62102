"""The pins of tests/pins.py catch the mistakes they are meant to catch.

Each case plants one plausible mistake in a copy of the oracle's source
(oracle/hd_oracle.c, text replacement on the CPU), builds that copy into a
temporary library, and checks that the pins named for it fail (an
AssertionError, or the oracle refusing to continue after a NaN).  The
unmodified oracle passes every pin in test_oracle_pins.py.
"""
import os
import subprocess

import pytest

from oracle import pyoracle
from tests import pins

SRC = os.path.join(os.path.dirname(pyoracle.__file__), "hd_oracle.c")

# (name, text in hd_oracle.c, replacement, pins that must fail)
MUTATIONS = [
    ("stress rows 4/3 and 2/3 swapped",
     "c2d3cv = (2. / 3.) * colvis, c4d3cv = (4. / 3.) * colvis;",
     "c2d3cv = (4. / 3.) * colvis, c4d3cv = (2. / 3.) * colvis;",
     ["pin_stress_rows", "pin_relaxation_waves"]),
    ("x-flux stress rows PXY and PXZ swapped",
     "F[PXY] = colvis * vy; F[PXZ] = colvis * vz; F[PYZ] = 0;",
     "F[PXY] = colvis * vz; F[PXZ] = colvis * vy; F[PYZ] = 0;",
     ["pin_stress_rows", "pin_relaxation_waves"]),
    ("y-flux stress rows PXY and PYZ swapped",
     "F[PXY] = colvis * vx; F[PXZ] = 0; F[PYZ] = colvis * vz;",
     "F[PXY] = colvis * vz; F[PXZ] = 0; F[PYZ] = colvis * vx;",
     ["pin_stress_rows", "pin_poiseuille"]),
    ("x-flux momentum rows take PXZ and PXY swapped",
     "F[MY] = M_y * vx + P_xy - Sxy;\n            F[MZ] = M_z * vx + P_xz - Sxz;",
     "F[MY] = M_y * vx + P_xz - Sxz;\n            F[MZ] = M_z * vx + P_xy - Sxy;",
     ["pin_stress_into_momentum", "pin_relaxation_waves"]),
    ("stress work dropped from the y energy flux",
     "F[EN] = (Ener + P) * vy + (P_yy - Syy) * vy + (P_xy - Sxy) * vx + (P_yz - Syz) * vz;",
     "F[EN] = (Ener + P) * vy + (P_yy - Syy) * vy + (P_yz - Syz) * vz;",
     ["pin_stress_into_momentum"]),
    ("implicit stress relaxation left out",
     "if (g->ivis == 1 && f >= PXX) v = faci * v;",
     "if (g->ivis == 1 && f >= PXX) v = v;",
     ["pin_relaxation_waves", "pin_poiseuille"]),
    ("low-side outflow decides on the minimum",
     "for (int p = 1; p < nface; p++) mxv = fmax(mxv, EX(p, normal));",
     "for (int p = 1; p < nface; p++) mxv = fmin(mxv, EX(p, normal));",
     ["pin_outflow_branches"]),
    ("high-side outflow decides on the maximum",
     "for (int p = 1; p < nface; p++) mnv = fmin(mnv, EX(p, normal));",
     "for (int p = 1; p < nface; p++) mnv = fmax(mnv, EX(p, normal));",
     ["pin_outflow_branches"]),
    ("x-high outflow mirrors the inflowing points",
     "if (EX(p, MX) > 0.0) EX(p, MX) = -IN(p, MX);",
     "if (EX(p, MX) < 0.0) EX(p, MX) = -IN(p, MX);",
     ["pin_outflow_branches"]),
    ("P2 derivative factor halved in the volume integral",
     "u1 = u1 + w[p] * 3. * c->bint[p][KX] * fx[p * nQ + f];",
     "u1 = u1 + w[p] * 1.5 * c->bint[p][KX] * fx[p * nQ + f];",
     ["pin_nb10_third_order"]),
    ("13-moment heat-flux row in the y flux",
     "F[T_QX] = 0; F[T_QY] = ct; F[T_QZ] = 0;",
     "F[T_QX] = ct; F[T_QY] = 0; F[T_QZ] = 0;",
     ["pin_13m_rotation", "pin_13m_waves"]),
    ("13-moment heat flux left out of the x energy flux",
     "F[EN] = (Ener + P) * vx + T_xx * vx + T_xy * vy + T_xz * vz + Q[T_QX];",
     "F[EN] = (Ener + P) * vx + T_xx * vx + T_xy * vy + T_xz * vz;",
     ["pin_13m_rotation", "pin_13m_waves"]),
    ("13-moment pzz with the wrong sign",
     "T_zz = -(T_xx + T_yy);",
     "T_zz = (T_xx + T_yy);",
     ["pin_13m_rotation", "pin_13m_waves"]),
    ("P2 normalisation halved",
     "c->cb[KXX] = c->cb[KYY] = c->cb[KZZ] = 5.;",
     "c->cb[KXX] = c->cb[KYY] = c->cb[KZZ] = 2.5;",
     ["pin_nb10_third_order"]),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(SRC).read()
    d = tmp_path_factory.mktemp("mutants")
    procs = []
    for n, (name, old, new, _) in enumerate(MUTATIONS):
        assert src.count(old) == 1, f"mutation site not unique or missing: {name}"
        path = d / f"m{n}.c"
        path.write_text(src.replace(old, new))
        lib = str(d / f"libm{n}.so")
        procs.append((lib, subprocess.Popen(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I",
                                             os.path.dirname(SRC), "-o", lib, str(path), "-lm"])))
    libs = []
    for lib, p in procs:
        assert p.wait() == 0
        libs.append(lib)
    return libs


@pytest.mark.parametrize("n", range(len(MUTATIONS)), ids=[m[0] for m in MUTATIONS])
def test_pins_fail_on_planted_mistake(mutant_libs, n):
    name, _, _, must_fail = MUTATIONS[n]
    L = pyoracle.load_library(mutant_libs[n])

    def Oracle(cfg):
        return pyoracle.Oracle(cfg, library=L)

    for pin in must_fail:
        with pytest.raises((AssertionError, pyoracle.OracleError)):
            getattr(pins, pin)(Oracle)
