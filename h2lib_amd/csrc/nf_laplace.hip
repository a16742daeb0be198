// Near-field kernels of the Laplace single and double layer potentials.
#include "nf_device.h"

namespace h2hip {
namespace nf {

int launch_laplace(int layer, int rb, int cb, uint32_t q, const KArgs &a,
                   uint32_t ntiles, const SPair *sp, const uint32_t *ncls,
                   hipStream_t stream) {
  return layer ? launch_family<K_LAP_DLP>(rb, cb, q, a, ntiles, sp, ncls,
                                          stream)
               : launch_family<K_LAP_SLP>(rb, cb, q, a, ntiles, sp, ncls,
                                          stream);
}

}  // namespace nf
}  // namespace h2hip
