"""Checkpoint files in the reference's text format (writeQ/readQ,
src/initialize.f90:470-565): round trips, rank layout and list-directed
parsing."""
import numpy as np

from hermeshd_amd.checkpoint import read_checkpoint, write_checkpoint


def _state(shape, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape)  # (nbasis, nQ, nz, ny, nx)


def test_round_trip_over_ranks(tmp_path):
    q = _state((4, 5, 4, 6, 8))
    names = write_checkpoint(str(tmp_path), "Qout", 7, q, 0.125, 3e-4, 12, mpi=(2, 3, 2))
    assert len(names) == 12 and names[0].endswith("Qout_p0000_d0007.dat")
    Q, t, dt, nout, mpi = read_checkpoint(str(tmp_path), "Qout", 7, (8, 6, 4), nQ=5, nbasis=4)
    assert (t, dt, nout, mpi) == (0.125, 3e-4, 12, (2, 3, 2))
    assert np.array_equal(Q, q.T)


def test_rank_numbering_is_row_major(tmp_path):
    """rank = (px*mpi_ny + py)*mpi_nz + pz; rank 1 is the second z block."""
    q = np.zeros((1, 1, 4, 2, 2))
    q[0, 0, 2:, 0:1, 0:1] = 5.0            # cells x<1, y<1, z>=2: block (0, 0, 1)
    write_checkpoint(str(tmp_path), "Qout", 0, q, 0.0, 0.0, 0, mpi=(2, 2, 2))
    lines = open(tmp_path / "Qout_p0001_d0000.dat").read().split("\n")
    assert [float(v) for v in lines[0].split()] == [5.0]


def test_reads_fortran_list_directed_text(tmp_path):
    """Values as gfortran writes them (padded, E and D exponents, r*value repeats)."""
    nx, ny, nz, nQ, nb = 3, 1, 1, 5, 1
    rows = ["   1.0000000000000000        2.5000000000000000       -3.0000000000000001E-003",
            "   2*1.0D0   4.0",
            "   0.0 0.0 0.0",
            "   7.5000000E-01   1.0000000   2.0000000",
            "  -1.0 -2.0 -3.0",
            "   0.50000000000000000        1.0000000000000000E-004           5           1           1           1"]
    (tmp_path / "Qout_p0000_d0003.dat").write_text("\n".join(rows) + "\n")
    Q, t, dt, nout, mpi = read_checkpoint(str(tmp_path), "Qout", 3, (nx, ny, nz), nQ=nQ, nbasis=nb)
    assert Q.shape == (3, 1, 1, 5, 1)
    assert list(Q[:, 0, 0, 0, 0]) == [1.0, 2.5, -3.0000000000000001e-3]
    assert list(Q[:, 0, 0, 1, 0]) == [1.0, 1.0, 4.0]
    assert list(Q[:, 0, 0, 3, 0]) == [0.75, 1.0, 2.0]
    assert (t, dt, nout, mpi) == (0.5, 1e-4, 5, (1, 1, 1))
