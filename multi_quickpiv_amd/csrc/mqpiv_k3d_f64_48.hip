// mqpiv_k3d_f64_48.hip -- instances of the 3-D pipeline: double transforms of length 48,
// 4 lane(s) per line.
#include "mqpiv_kernels3d.hpp"

namespace mqpiv {
template size_t scratch3_per_ia<48, double, 4>(int, int, int, int);
template hipError_t launch3d_chunk<48, double, 4>(const Job&, void*, int64_t, int64_t, int64_t, const double*,
                                                   hipStream_t);
}  // namespace mqpiv
