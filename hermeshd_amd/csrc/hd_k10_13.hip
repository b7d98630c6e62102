// hd_k10_13.hip -- kernels for nbasis 10 (quadratic), iquad 3, the 13-moment model.
#include "hd_impl.h"
#include "hd_impl_vjp.h"
namespace hd {
bool hd_pick_10_13(int rm, Ops* o) { return pick_rm<10, 3, 13>(rm, o); }
}  // namespace hd
