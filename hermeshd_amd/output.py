"""Field output of output_vtk (src/output.f90:14-318): the derived fields are
sampled on the device (``hd_sample``) and written as a VTK XML rectilinear
grid (``.vtr``) by the small writer below (inline base64 "binary" arrays),
so no VTK library is needed.

File names follow the reference: ``<outdir>/perseus_p<rank>_t<nout>.vtr``
with 4-digit rank and output index.  Values are in the reference's physical
units: density times n0, velocity times v0, temperature P/rho times
te0/eV_per_K (Kelvin), vorticity over t0 (src/output.f90:123-300).
"""
from __future__ import annotations

import base64
import ctypes as C
import os
import struct

import numpy as np
import torch

from . import _lib
from .config import HDConfig

# field bits of hd_sample (HD_OUT_*), with the reference's VTK names
FIELDS = ["Density", "Log density", "vx", "vy", "vz", "Temperature", "Internal energy density",
          "Total energy density", "Entropy density", "Pressure", "pxx", "pyy", "pzz", "Vorticity"]
BIT = {name: i for i, name in enumerate(FIELDS)}
# the reference's output switches (src/input.f90:87-96) -> field bits
SWITCHES = {
    "nstdout": ["Density"], "nstldout": ["Log density"], "nstvout": ["vx", "vy", "vz"],
    "nsttout": ["Temperature"], "nsteiout": ["Internal energy density"], "nstenout": ["Total energy density"],
    "nstesout": ["Entropy density"], "nstpout": ["Pressure"], "nststout": ["pxx", "pyy", "pzz"],
    "nstvrout": ["Vorticity"],
}
# the reference's defaults: density, velocity, temperature, pressure
DEFAULT_SWITCHES = ("nstdout", "nstvout", "nsttout", "nstpout")


def reference_units(mu: float = 39.948):
    """(n0, v0, te0/eV_per_K, t0) of src/params.f90:80-88 with the shipped mu."""
    L0, t0, n0 = 1.0, 1.0, 2.5e25
    v0 = L0 / t0
    p0 = mu * 1.67e-27 * n0 * v0 ** 2
    te0 = p0 / n0 / 1.6e-19
    return n0, v0, te0 / 8.61728e-5, t0


def unit_scales(units=None) -> np.ndarray:
    n0, v0, tK, t0 = reference_units() if units is None else units
    sc = np.ones(len(FIELDS))
    sc[BIT["Density"]] = n0
    sc[BIT["Log density"]] = n0
    sc[[BIT["vx"], BIT["vy"], BIT["vz"]]] = v0
    sc[BIT["Temperature"]] = tK
    sc[BIT["Vorticity"]] = 1.0 / t0
    return sc


def field_mask(names) -> int:
    m = 0
    for n in names:
        m |= 1 << BIT[n]
    return m


def sample(solver, q: torch.Tensor, names, nvtk: int = 1, scales=None) -> dict:
    """The named fields of q as float32 device tensors (nz*nvtk, ny*nvtk,
    nx*nvtk), in code units unless scales (one factor per field) are given."""
    cfg = solver.cfg
    names = [n for n in FIELDS if n in set(names)]
    mask = field_mask(names)
    out = torch.empty((len(names), cfg.nz * nvtk, cfg.ny * nvtk, cfg.nx * nvtk), dtype=torch.float32,
                      device=solver.device)
    sc = None
    if scales is not None:
        sc = np.ascontiguousarray(scales, dtype=np.float64)
    _lib.check(_lib.lib().hd_sample(solver._h, solver._ptr(q), int(nvtk), mask,
                                    None if sc is None else sc.ctypes.data_as(C.POINTER(C.c_double)),
                                    C.c_void_p(out.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return {n: out[i] for i, n in enumerate(names)}


def _b64(a: np.ndarray) -> str:
    raw = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")


def write_vtr(path: str, cfg: HDConfig, fields: dict, nvtk: int = 1, origin=None) -> None:
    """VTK XML RectilinearGrid with cell data.  fields: name -> array
    (nz*nvtk, ny*nvtk, nx*nvtk) (scalar) or (3, ...) (vector).  The cell edges
    are the reference's x_xml_rect = xvtk(i) - dxvtk/2 (src/output.f90:61-69)."""
    nnx, nny, nnz = cfg.nx * nvtk, cfg.ny * nvtk, cfg.nz * nvtk
    lo = origin if origin is not None else (-cfg.lx / 2.0, -cfg.ly / 2.0, -cfg.lz / 2.0)
    edges = []
    for n, L, o in ((nnx, cfg.lx, lo[0]), (nny, cfg.ly, lo[1]), (nnz, cfg.lz, lo[2])):
        d = L / n
        edges.append((o + np.arange(n + 1) * d).astype(np.float32))
    ext = f"0 {nnx} 0 {nny} 0 {nnz}"
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">',
             f'  <RectilinearGrid WholeExtent="{ext}">', f'    <Piece Extent="{ext}">', '      <CellData>']
    for name, a in fields.items():
        a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
        a = a.astype(np.float32)
        if a.ndim == 4:  # vector: interleave components, x fastest over cells
            comp = a.shape[0]
            data = np.stack([a[c].ravel() for c in range(comp)], axis=1)
            lines.append(f'        <DataArray type="Float32" Name="{name}" NumberOfComponents="{comp}" '
                         f'format="binary">{_b64(data)}</DataArray>')
        else:
            lines.append(f'        <DataArray type="Float32" Name="{name}" format="binary">{_b64(a.ravel())}'
                         '</DataArray>')
    lines += ['      </CellData>', '      <Coordinates>']
    for ax, e in zip("XYZ", edges):
        lines.append(f'        <DataArray type="Float32" Name="{ax}" format="binary">{_b64(e)}</DataArray>')
    lines += ['      </Coordinates>', '    </Piece>', '  </RectilinearGrid>', '</VTKFile>', '']
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


def read_vtr(path: str) -> dict:
    """Cell arrays of a file written by write_vtr: name -> float32 array (flat,
    or (ncell, ncomp) for vectors), plus the X/Y/Z edges."""
    import re
    text = open(path).read()
    out = {}
    for m in re.finditer(r'<DataArray type="Float32" Name="([^"]+)"( NumberOfComponents="(\d+)")? '
                         r'format="binary">([^<]*)</DataArray>', text):
        raw = base64.b64decode(m.group(4))
        n = struct.unpack("<I", raw[:4])[0]
        a = np.frombuffer(raw[4:4 + n], dtype=np.float32)
        if m.group(3):
            a = a.reshape(-1, int(m.group(3)))
        out[m.group(1)] = a
    return out


def output_vtk(solver, q: torch.Tensor, nout: int, iam: int = 0, outdir: str = "output", switches=DEFAULT_SWITCHES,
               nvtk: int = 1, units=None) -> str:
    """output_vtk (src/output.f90:14-318): the fields of the given switches
    (names of src/input.f90:87-96) in physical units, to
    <outdir>/perseus_p<iam>_t<nout>.vtr.  Returns the file name."""
    names = []
    for sw in switches:
        names += SWITCHES[sw]
    vals = sample(solver, q, names, nvtk=nvtk, scales=unit_scales(units))
    fields = {}
    for sw in switches:
        if sw == "nstvout":
            fields["Velocity"] = torch.stack([vals["vx"], vals["vy"], vals["vz"]])
        elif sw == "nststout":
            fields["Isotropic stress"] = torch.stack([vals["pxx"], vals["pyy"], vals["pzz"]])
        else:
            fields[SWITCHES[sw][0]] = vals[SWITCHES[sw][0]]
    name = os.path.join(outdir, f"perseus_p{iam:04d}_t{nout:04d}.vtr")
    write_vtr(name, solver.cfg, fields, nvtk=nvtk)
    return name
