// hd_k4_vjp.hip -- the stage transposes for nbasis 4 (linear), iquad 2,
// instantiated apart from the forward kernels of hd_k4.hip.
#include "hd_impl_vjp.h"
namespace hd {
#define HD_VJP_INST(NQ, RM)                                                                                     \
    template void Impl<4, 2, NQ, RM>::stage_vjp(hd_ctx*, const double*, const double*, double*,                  \
                                                const adj::AdjParams&, hipStream_t);
HD_VJP_INST(5, RIEMANN_CENTRAL)
HD_VJP_INST(5, RIEMANN_HLLC)
HD_VJP_INST(5, RIEMANN_ROE)
HD_VJP_INST(11, RIEMANN_CENTRAL)
HD_VJP_INST(11, RIEMANN_HLLC)
HD_VJP_INST(11, RIEMANN_ROE)
HD_VJP_INST(13, RIEMANN_CENTRAL)
HD_VJP_INST(13, RIEMANN_HLLC)
HD_VJP_INST(13, RIEMANN_ROE)
#undef HD_VJP_INST
}  // namespace hd
