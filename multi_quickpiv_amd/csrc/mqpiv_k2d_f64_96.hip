// mqpiv_k2d_f64_96.hip -- instances of the fused 2-D kernel:
// double transforms of length 96 (lanes per line: mqpiv_lpl2d.h).
#include "mqpiv_kernels2d.hpp"
#include "mqpiv_lpl2d.h"

namespace mqpiv {
template size_t lds_bytes_2d<96, double, Lpl2<double, 96>::v>(int, int, int, int);
template hipError_t launch_2d_l<96, double, Lpl2<double, 96>::v>(int, const Job&, const Job*, const int64_t*, int, int64_t,
                                                int, int, int, char*, hipStream_t, int32_t*);
template hipError_t launch_2d_refine<96, double, Lpl2<double, 96>::v>(int, const Job&, const Job*, const int64_t*,
                                                     int, int64_t, int, int, int, const int32_t*, char*,
                                                     hipStream_t);
}  // namespace mqpiv
