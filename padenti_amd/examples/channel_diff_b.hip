// Instantiations of ChannelDiff (channel_diff.hpp), part 2 of 2.
#include "channel_diff.hpp"

using T_cd_u16x2_u32 = ChannelDiff<uint16_t, 2, uint32_t>;
PD_DEFINE_FEATURE(cd_u16x2_u32, T_cd_u16x2_u32)
using T_cd_i8x7_u16 = ChannelDiff<int8_t, 7, uint16_t>;
PD_DEFINE_FEATURE(cd_i8x7_u16, T_cd_i8x7_u16)
using T_cd_u16x16_u64 = ChannelDiff<uint16_t, 16, uint64_t>;
PD_DEFINE_FEATURE(cd_u16x16_u64, T_cd_u16x16_u64)
using T_cd_f64x4_f32 = ChannelDiff<double, 4, float>;
PD_DEFINE_FEATURE(cd_f64x4_f32, T_cd_f64x4_f32)
using T_cd_i32x3_i16 = ChannelDiff<int32_t, 3, int16_t>;
PD_DEFINE_FEATURE(cd_i32x3_i16, T_cd_i32x3_i16)
