// Shape of the fused 2-D kernel for each (scalar type, transform length):
// Lpl2 = lanes that share one line of the transform (1, 2, 4 or 8).  Tuning
// builds may override an entry with -DMQPIV_LPL_<F32|F64>_<L>=n.
#pragma once

namespace mqpiv {

template <typename T, int L>
struct Lpl2;
#ifndef MQPIV_LPL_F32_16
#define MQPIV_LPL_F32_16 1
#endif
template <> struct Lpl2<float, 16> { static constexpr int v = MQPIV_LPL_F32_16; };
#ifndef MQPIV_LPL_F32_32
#define MQPIV_LPL_F32_32 2
#endif
template <> struct Lpl2<float, 32> { static constexpr int v = MQPIV_LPL_F32_32; };
#ifndef MQPIV_LPL_F32_48
#define MQPIV_LPL_F32_48 4
#endif
template <> struct Lpl2<float, 48> { static constexpr int v = MQPIV_LPL_F32_48; };
#ifndef MQPIV_LPL_F32_64
#define MQPIV_LPL_F32_64 4
#endif
template <> struct Lpl2<float, 64> { static constexpr int v = MQPIV_LPL_F32_64; };
#ifndef MQPIV_LPL_F32_96
#define MQPIV_LPL_F32_96 8
#endif
template <> struct Lpl2<float, 96> { static constexpr int v = MQPIV_LPL_F32_96; };
#ifndef MQPIV_LPL_F32_128
#define MQPIV_LPL_F32_128 4
#endif
template <> struct Lpl2<float, 128> { static constexpr int v = MQPIV_LPL_F32_128; };
#ifndef MQPIV_LPL_F64_16
#define MQPIV_LPL_F64_16 1
#endif
template <> struct Lpl2<double, 16> { static constexpr int v = MQPIV_LPL_F64_16; };
#ifndef MQPIV_LPL_F64_32
#define MQPIV_LPL_F64_32 2
#endif
template <> struct Lpl2<double, 32> { static constexpr int v = MQPIV_LPL_F64_32; };
#ifndef MQPIV_LPL_F64_48
#define MQPIV_LPL_F64_48 4
#endif
template <> struct Lpl2<double, 48> { static constexpr int v = MQPIV_LPL_F64_48; };
#ifndef MQPIV_LPL_F64_64
#define MQPIV_LPL_F64_64 4
#endif
template <> struct Lpl2<double, 64> { static constexpr int v = MQPIV_LPL_F64_64; };
#ifndef MQPIV_LPL_F64_96
#define MQPIV_LPL_F64_96 4
#endif
template <> struct Lpl2<double, 96> { static constexpr int v = MQPIV_LPL_F64_96; };
#ifndef MQPIV_LPL_F64_128
#define MQPIV_LPL_F64_128 4
#endif
template <> struct Lpl2<double, 128> { static constexpr int v = MQPIV_LPL_F64_128; };

}  // namespace mqpiv
