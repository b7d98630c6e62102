"""Checkpoints in the reference's text format, so that runs move between the
reference and this build.

The reference writes one file per MPI rank, ``data/<prefix>_p<rank>_d<dump>.dat``
(rank and dump as 4 digits), each holding its subdomain Q(nx,ny,nz,nQ,nbasis)
as one list-directed line of nx values per (mode, field, k, j), modes
outermost, followed by a line ``t dt nout mpi_nx mpi_ny mpi_nz``
(writeQ/readQ, src/initialize.f90:470-565).  Ranks are numbered row-major
over the (mpi_nx, mpi_ny, mpi_nz) Cartesian grid (MPI_CART_CREATE,
src/initialize.f90:250-270), so rank = (px*mpi_ny + py)*mpi_nz + pz.

``write_checkpoint`` splits a whole-grid state into such files;
``read_checkpoint`` reads the files of one dump and stitches the full grid.
Host-side I/O only: the device state is copied once per call.
"""
from __future__ import annotations

import os

import numpy as np


def _fname(directory: str, prefix: str, rank: int, dump: int) -> str:
    if len(prefix) != 4:
        raise ValueError("the reference's file prefix has 4 characters")
    return os.path.join(directory, f"{prefix}_p{rank:04d}_d{dump:04d}.dat")


def _as_fortran(q) -> np.ndarray:
    """(nx,ny,nz,nQ,nbasis) float64 view of a state given as a torch/numpy
    (nbasis,nQ,nz,ny,nx) C-ordered array or an F-ordered (nx,ny,nz,nQ,nbasis) one."""
    try:
        import torch
        if isinstance(q, torch.Tensor):
            q = q.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    q = np.asarray(q, dtype=np.float64)
    if q.flags["C_CONTIGUOUS"] and not q.flags["F_CONTIGUOUS"]:
        return q.T  # (nbasis,nQ,nz,ny,nx) C order == (nx,ny,nz,nQ,nbasis) F order
    return q


def write_checkpoint(directory: str, prefix: str, dump: int, q, t: float, dt: float, nout: int,
                     mpi=(1, 1, 1)) -> list[str]:
    """Write the whole-grid state q as the per-rank files of an
    mpi = (mpi_nx, mpi_ny, mpi_nz) decomposition.  Returns the file names."""
    Q = _as_fortran(q)
    NX, NY, NZ, nQ, nb = Q.shape
    px, py, pz = (int(v) for v in mpi)
    if NX % px or NY % py or NZ % pz:
        raise ValueError("the grid must divide evenly over the ranks")
    nx, ny, nz = NX // px, NY // py, NZ // pz
    os.makedirs(directory, exist_ok=True)
    names = []
    for cx in range(px):
        for cy in range(py):
            for cz in range(pz):
                rank = (cx * py + cy) * pz + cz
                sub = Q[cx * nx:(cx + 1) * nx, cy * ny:(cy + 1) * ny, cz * nz:(cz + 1) * nz]
                name = _fname(directory, prefix, rank, dump)
                with open(name, "w") as fh:
                    for r in range(nb):
                        for f in range(nQ):
                            for k in range(nz):
                                for j in range(ny):
                                    fh.write(" " + " ".join(repr(float(v)) for v in sub[:, j, k, f, r]) + "\n")
                    fh.write(f" {float(t)!r} {float(dt)!r} {int(nout)} {px} {py} {pz}\n")
                names.append(name)
    return names


def _numbers(line: str) -> list[str]:
    # list-directed output may use D exponents and r*value repeats
    out = []
    for tok in line.replace(",", " ").split():
        tok = tok.replace("D", "E").replace("d", "e")
        if "*" in tok:
            n, v = tok.split("*", 1)
            out.extend([v] * int(n))
        else:
            out.append(tok)
    return out


def read_checkpoint(directory: str, prefix: str, dump: int, grid_shape, nQ: int, nbasis: int):
    """Read dump `dump` of a run on a grid_shape = (NX,NY,NZ) grid.  The
    decomposition (mpi_nx, mpi_ny, mpi_nz) is read from rank 0's last line.
    Returns (Q as an F-ordered (NX,NY,NZ,nQ,nbasis) array, t, dt, nout,
    (mpi_nx, mpi_ny, mpi_nz))."""
    with open(_fname(directory, prefix, 0, dump)) as fh:
        last = [ln for ln in fh.read().splitlines() if ln.strip()][-1]
    tail0 = _numbers(last)
    if len(tail0) < 6:
        raise ValueError("rank 0's file does not end with the t, dt, nout, mpi_nx, mpi_ny, mpi_nz line")
    dims = tuple(int(float(v)) for v in tail0[-3:])
    NX, NY, NZ = (int(v) for v in grid_shape)
    if NX % dims[0] or NY % dims[1] or NZ % dims[2]:
        raise ValueError(f"grid {grid_shape} does not divide over the file's ranks {dims}")
    nx, ny, nz = NX // dims[0], NY // dims[1], NZ // dims[2]

    def read_one(rank):
        name = _fname(directory, prefix, rank, dump)
        with open(name) as fh:
            lines = fh.read().splitlines()
        nrows = nbasis * nQ * nz * ny
        body = []
        for line in lines:
            body.extend(_numbers(line))
        data = np.array([float(v) for v in body[:nrows * nx]], dtype=np.float64)
        if data.size != nrows * nx:
            raise ValueError(f"{name}: expected {nrows * nx} values, found {data.size}")
        tail = body[nrows * nx:]
        if len(tail) < 6:
            raise ValueError(f"{name}: missing the t, dt, nout, mpi_nx, mpi_ny, mpi_nz line")
        sub = data.reshape(nbasis, nQ, nz, ny, nx).T  # -> (nx,ny,nz,nQ,nbasis)
        meta = (float(tail[0]), float(tail[1]), int(float(tail[2])), int(float(tail[3])), int(float(tail[4])),
                int(float(tail[5])))
        return sub, meta

    sub0, meta0 = read_one(0)
    t, dt, nout, px, py, pz = meta0
    Q = np.zeros((nx * px, ny * py, nz * pz, nQ, nbasis), dtype=np.float64, order="F")
    for cx in range(px):
        for cy in range(py):
            for cz in range(pz):
                rank = (cx * py + cy) * pz + cz
                sub, meta = (sub0, meta0) if rank == 0 else read_one(rank)
                if meta[3:] != (px, py, pz):
                    raise ValueError(f"rank {rank}: decomposition {meta[3:]} differs from rank 0's {(px, py, pz)}")
                Q[cx * nx:(cx + 1) * nx, cy * ny:(cy + 1) * ny, cz * nz:(cz + 1) * nz] = sub
    return Q, t, dt, nout, (px, py, pz)
