// mqpiv_k2d_f32_64.hip -- instances of the fused 2-D kernel:
// float transforms of length 64 (lanes per line: mqpiv_lpl2d.h).
#include "mqpiv_kernels2d.hpp"
#include "mqpiv_lpl2d.h"

namespace mqpiv {
template size_t lds_bytes_2d<64, float, Lpl2<float, 64>::v>(int, int, int, int);
template hipError_t launch_2d_l<64, float, Lpl2<float, 64>::v>(int, const Job&, const Job*, const int64_t*, int, int64_t,
                                                int, int, int, char*, hipStream_t, int32_t*);
}  // namespace mqpiv
