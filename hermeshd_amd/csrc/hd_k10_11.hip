// hd_k10_11.hip -- kernels for nbasis 10 (quadratic), iquad 3, eleven fields.
#include "hd_impl.h"
#include "hd_impl_vjp.h"
namespace hd {
bool hd_pick_10_11(int rm, Ops* o) { return pick_rm<10, 3, 11>(rm, o); }
}  // namespace hd
