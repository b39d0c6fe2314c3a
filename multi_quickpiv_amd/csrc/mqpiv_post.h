// Launchers of the device vector-field filters (mqpiv_post.hip), shared with
// the C-ABI translation unit.  dt is the element type: MQPIV_F32 (0) or
// MQPIV_F64 (1).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mqpiv {
namespace post {

// Vector grid of a field (N components per vector, column-major grid) and the
// clipped box radius of a stencil.
struct VGrid {
  int32_t nd;
  int32_t n[3];  // vectors per dim (unused dims 1)
  int32_t r[3];  // box radius per dim (unused dims 0)
  int64_t nv;
};

}  // namespace post

hipError_t pp_magnitudes(int dt, const post::VGrid& g, const void* VF, void* M, hipStream_t s);
hipError_t pp_std_mask(int dt, int64_t n, const void* x, double nsd, uint8_t* mask, double* scratch,
                       hipStream_t s);
size_t pp_std_scratch_bytes();
hipError_t pp_th_mask(int dt, int64_t n, const void* x, double th, int cmp, uint8_t* mask, hipStream_t s);
hipError_t pp_replace(int dt, const post::VGrid& g, const void* VF, const uint8_t* mask, int how, void* out,
                      hipStream_t s);
hipError_t pp_average(int dt, const post::VGrid& g, const void* VF, void* out, hipStream_t s);
hipError_t pp_collectiveness(int dt, const post::VGrid& g, const void* VF, void* M, double min_dot, float* out,
                             hipStream_t s);
hipError_t pp_simavg(int dt, const post::VGrid& g, const void* VF, void* M, double min_dot, int normalize,
                     int scale, int expo, void* out, hipStream_t s);
hipError_t pp_lmm(int dt, const post::VGrid& g, const void* VF, void* M, double ratio, void* out,
                  uint8_t* removed, hipStream_t s);

}  // namespace mqpiv
