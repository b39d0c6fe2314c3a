// mqpiv_k3d_f32_32.hip -- instances of the 3-D pipeline: float transforms of length 32,
// 2 lane(s) per line.
#include "mqpiv_kernels3d.hpp"

namespace mqpiv {
template size_t scratch3_per_ia<32, float, 2>(int, int, int, int);
template hipError_t launch3d_chunk<32, float, 2>(const Job&, void*, int64_t, int64_t, int64_t, const double*,
                                                   hipStream_t);
}  // namespace mqpiv
