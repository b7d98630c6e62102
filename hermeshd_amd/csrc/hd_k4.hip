// hd_k4.hip -- kernels for nbasis 4 (linear), iquad 2.
#include "hd_impl.h"
namespace hd {
bool hd_pick_4(int nq, int rm, Ops* o) {
    if (nq == 5) return pick_rm<4, 2, 5>(rm, o);
    if (nq == 11) return pick_rm<4, 2, 11>(rm, o);
    if (nq == 13) return pick_rm<4, 2, 13>(rm, o);
    return false;
}
}  // namespace hd
