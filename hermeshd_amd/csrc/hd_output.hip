// hd_output.hip -- derived-field sampling of output_vtk (src/output.f90:14-318)
// on the device: every cell is sampled at nvtk^3 interior points
// (set_vtk_vals_3D, src/basis_funcs.f90:90-200) and the selected fields are
// written as float32 planes, ready for a VTK writer on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "hd_ctx.h"

namespace hd {
namespace out {

enum : int {
    F_DENSITY = 0, F_LOGDENSITY = 1, F_VX = 2, F_VY = 3, F_VZ = 4, F_TEMPERATURE = 5, F_EINT = 6, F_ETOT = 7,
    F_ENTROPY = 8, F_PRESSURE = 9, F_PXX = 10, F_PYY = 11, F_PZZ = 12, F_VORTICITY = 13, NFIELDS = 14
};

struct SampleParams {
    int nvtk, nb, nQ, nsel;
    int sel[NFIELDS];      // selected fields in output order
    double scale[NFIELDS]; // unit factors (log density: applied inside the log)
    double gm1, dxi;
};

__device__ __forceinline__ double leg(int a, double x) { return a == 0 ? 1.0 : a == 1 ? x : 1.5 * x * x - 0.5; }
__device__ __forceinline__ double dleg(int a, double x) { return a == 0 ? 0.0 : a == 1 ? 1.0 : 3.0 * x; }

// Legendre degrees (x, y, z) of the modes 1, x, y, z, yz, zx, xy, (xyz | P2x, P2y, P2z)
__device__ __forceinline__ void mode_degrees(int nb, int r, int& a, int& b, int& c) {
    const int d4[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int d7[3][3] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
    if (r < 4) { a = d4[r][0]; b = d4[r][1]; c = d4[r][2]; return; }
    if (r < 7) { a = d7[r - 4][0]; b = d7[r - 4][1]; c = d7[r - 4][2]; return; }
    if (nb == 8) { a = b = c = 1; return; }
    a = r == 7 ? 2 : 0; b = r == 8 ? 2 : 0; c = r == 9 ? 2 : 0;
}

// sub-cell point coordinate of set_vtk_vals_3D: (2i - 1 - nvtk)/nvtk, i = 1..nvtk
__device__ __forceinline__ double sub_point(int i0, int nvtk) { return (double)(2 * i0 + 1 - nvtk) / nvtk; }

__global__ __launch_bounds__(256) void k_sample(const double* __restrict__ q, float* __restrict__ out, Grid g,
                                                SampleParams sp) {
    const int nnx = g.nx * sp.nvtk, nny = g.ny * sp.nvtk, nnz = g.nz * sp.nvtk;
    const long npts = (long)nnx * nny * nnz;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= npts) return;
    const int i = (int)(id % nnx), j = (int)((id / nnx) % nny), k = (int)(id / ((long)nnx * nny));
    const int ir = i / sp.nvtk, jr = j / sp.nvtk, kr = k / sp.nvtk;
    const double x = sub_point(i % sp.nvtk, sp.nvtk), y = sub_point(j % sp.nvtk, sp.nvtk),
                 z = sub_point(k % sp.nvtk, sp.nvtk);
    const long c = (long)ir + (long)g.nx * ((long)jr + (long)g.ny * kr);
    double v[11], dx[3] = {0., 0., 0.}, dy[3] = {0., 0., 0.};  // dx/dy of rh, mx, my
    for (int f = 0; f < sp.nQ; f++) v[f] = 0.;
    for (int r = 0; r < sp.nb; r++) {
        int a, b, cc;
        mode_degrees(sp.nb, r, a, b, cc);
        const double phi = leg(a, x) * leg(b, y) * leg(cc, z);
        const double px = dleg(a, x) * leg(b, y) * leg(cc, z);
        const double py = leg(a, x) * dleg(b, y) * leg(cc, z);
        for (int f = 0; f < sp.nQ; f++) v[f] += phi * q[((long)r * sp.nQ + f) * g.N + c];
        for (int f = 0; f < 3; f++) {
            const double m = q[((long)r * sp.nQ + f) * g.N + c];
            dx[f] += px * m;
            dy[f] += py * m;
        }
    }
    const double dn = v[RH], dni = 1. / dn;
    const double vx = v[MX] * dni, vy = v[MY] * dni, vz = v[MZ] * dni;
    const double U = v[EN] - 0.5 * dn * (vx * vx + vy * vy + vz * vz);
    const double P = sp.gm1 * U;
    for (int s = 0; s < sp.nsel; s++) {
        const int f = sp.sel[s];
        double val = 0.;
        switch (f) {
            case F_DENSITY: val = dn; break;
            case F_LOGDENSITY: val = log(dn * sp.scale[f]) / log(10.); break;
            case F_VX: val = vx; break;
            case F_VY: val = vy; break;
            case F_VZ: val = vz; break;
            case F_TEMPERATURE: val = P * dni; break;
            case F_EINT: val = U; break;
            case F_ETOT: val = v[EN]; break;
            case F_ENTROPY: val = P * pow(dni, sp.gm1); break;
            case F_PRESSURE: val = P; break;
            case F_PXX: val = sp.nQ > PXX ? v[PXX] : 0.; break;
            case F_PYY: val = sp.nQ > PYY ? v[PYY] : 0.; break;
            case F_PZZ: val = sp.nQ > PZZ ? v[PZZ] : 0.; break;
            case F_VORTICITY: {
                // -2*(dv/dx - du/dy) in cell coordinates, times dxi, as output_vtk writes it
                const double dxvy = (dn * dx[2] - v[MY] * dx[0]) / (dn * dn);
                const double dyvx = (dn * dy[1] - v[MX] * dy[0]) / (dn * dn);
                val = -2. * (dxvy - dyvx) * sp.dxi;
                break;
            }
        }
        if (f != F_LOGDENSITY) val *= sp.scale[f];
        out[(long)s * npts + id] = (float)val;
    }
}

}  // namespace out
}  // namespace hd

using namespace hd;

extern "C" int hd_sample(hd_ctx* c, const double* q, int32_t nvtk, uint32_t fields, const double* scales,
                         float* out, void* stream) {
    if (!c || !q || !out) return fail(HD_ERR_INVALID, "NULL argument");
    if (nvtk < 1 || nvtk > 4) return fail(HD_ERR_INVALID, "nvtk must be 1..4 (set_vtk_vals_3D)");
    out::SampleParams sp;
    sp.nvtk = nvtk; sp.nb = c->nb; sp.nQ = c->nQ; sp.nsel = 0;
    for (int f = 0; f < out::NFIELDS; f++) {
        sp.scale[f] = scales ? scales[f] : 1.0;
        if (fields & (1u << f)) sp.sel[sp.nsel++] = f;
    }
    if (fields >> out::NFIELDS) return fail(HD_ERR_INVALID, "unknown field bit");
    if (sp.nsel == 0) return HD_OK;
    sp.gm1 = c->cfg.gamma - 1.0;
    sp.dxi = c->grid.dxi;
    const long npts = c->N * (long)nvtk * nvtk * nvtk;
    hipLaunchKernelGGL(out::k_sample, dim3(blocks_for(npts, 256)), dim3(256), 0, (hipStream_t)stream, q, out, c->grid,
                       sp);
    HIPCHK(hipGetLastError());
    return HD_OK;
}
