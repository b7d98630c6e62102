// hd_k10_5.hip -- kernels for nbasis 10 (quadratic), iquad 3, five fields.
#include "hd_impl.h"
#include "hd_impl_vjp.h"
namespace hd {
bool hd_pick_10_5(int rm, Ops* o) { return pick_rm<10, 3, 5>(rm, o); }
}  // namespace hd
