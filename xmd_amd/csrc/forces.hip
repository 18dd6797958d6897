// forces.hip -- translation unit of the force kernel, built with
// -ffp-contract=fast (see force_kernel.hpp).
#include "force_kernel.hpp"

namespace xmd_k {

void launch_force(bool mixed, bool eval, int nslot, hipStream_t stream, const Dev &D, const Derived &G,
                  const TermTab<float> &Tf, const TermTab<double> &Td, int p) {
    const int grid = nslot * FORCE_LPB / BLK;
    if (mixed) {
        if (eval) k_force<float, true, FORCE_LPB><<<grid, BLK, 0, stream>>>(D, G, Tf, p);
        else k_force<float, false, FORCE_LPB><<<grid, BLK, 0, stream>>>(D, G, Tf, p);
    } else {
        if (eval) k_force<double, true, FORCE_LPB><<<grid, BLK, 0, stream>>>(D, G, Td, p);
        else k_force<double, false, FORCE_LPB><<<grid, BLK, 0, stream>>>(D, G, Td, p);
    }
}

}  // namespace xmd_k
