// forces.hip -- translation unit of the force kernel, built with
// -ffp-contract=fast-honor-pragmas (see force_kernel.hpp).
#include "force_kernel.hpp"

namespace xmd_k {

template <int LPB>
static void launch_lpb(bool mixed, bool eval, int nslot, hipStream_t stream, const Dev &D, const Derived &G,
                       const TermTab<float> &Tf, const TermTab<double> &Td, int p, int mutate) {
    const int grid = nslot * LPB / FBLK;
    if (mixed) {
        if (eval) k_force<float, true, LPB><<<grid, FBLK, 0, stream>>>(D, G, Tf, p, mutate);
        else k_force<float, false, LPB><<<grid, FBLK, 0, stream>>>(D, G, Tf, p, mutate);
    } else {
        if (eval) k_force<double, true, LPB><<<grid, FBLK, 0, stream>>>(D, G, Td, p, mutate);
        else k_force<double, false, LPB><<<grid, FBLK, 0, stream>>>(D, G, Td, p, mutate);
    }
}

void launch_force(bool mixed, bool eval, int lpb, int nslot, hipStream_t stream, const Dev &D, const Derived &G,
                  const TermTab<float> &Tf, const TermTab<double> &Td, int p, int mutate) {
    if (lpb == 4) launch_lpb<4>(mixed, eval, nslot, stream, D, G, Tf, Td, p, mutate);
    else if (lpb == 16) launch_lpb<16>(mixed, eval, nslot, stream, D, G, Tf, Td, p, mutate);
    else launch_lpb<8>(mixed, eval, nslot, stream, D, G, Tf, Td, p, mutate);
}

}  // namespace xmd_k
