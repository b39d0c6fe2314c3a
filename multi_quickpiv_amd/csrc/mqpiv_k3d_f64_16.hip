// mqpiv_k3d_f64_16.hip -- instances of the 3-D pipeline: double transforms of length 16,
// 2 lane(s) per line.
#include "mqpiv_kernels3d.hpp"

namespace mqpiv {
template size_t scratch3_per_ia<16, double, 2>(int, int, int, int);
template hipError_t launch3d_chunk<16, double, 2>(const Job&, void*, int64_t, int64_t, int64_t, const double*,
                                                   hipStream_t);
}  // namespace mqpiv
