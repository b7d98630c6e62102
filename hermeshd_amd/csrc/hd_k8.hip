// hd_k8.hip -- kernels for nbasis 8 (trilinear), iquad 2.
#include "hd_impl.h"
#include "hd_impl_vjp.h"
namespace hd {
bool hd_pick_8(int nq, int rm, Ops* o) {
    if (nq == 5) return pick_rm<8, 2, 5>(rm, o);
    if (nq == 11) return pick_rm<8, 2, 11>(rm, o);
    return false;
}
}  // namespace hd
