// mqpiv_kernels3d.hip -- 3-D cross-correlation PIV (placeholder until the 3-D
// kernels land: every 3-D request is reported as unsupported).
#include <hip/hip_runtime.h>
#include "mqpiv_internal.h"

namespace mqpiv {
int fft_len_3d(int) { return -1; }
size_t scratch_3d_bytes(const Job&, int64_t) { return 0; }
hipError_t launch_piv3d(const Job&, void*, size_t, hipStream_t) { return hipErrorInvalidValue; }
}  // namespace mqpiv
