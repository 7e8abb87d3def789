// Instantiations of ChannelDiff (channel_diff.hpp), part 1 of 2.
#include "channel_diff.hpp"

using T_cd_u8x10_f64 = ChannelDiff<uint8_t, 10, double>;
PD_DEFINE_FEATURE(cd_u8x10_f64, T_cd_u8x10_f64)
using T_cd_i16x5_i32 = ChannelDiff<int16_t, 5, int32_t>;
PD_DEFINE_FEATURE(cd_i16x5_i32, T_cd_i16x5_i32)
using T_cd_u16x3_u8 = ChannelDiff<uint16_t, 3, uint8_t>;
PD_DEFINE_FEATURE(cd_u16x3_u8, T_cd_u16x3_u8)
using T_cd_f32x6_i64 = ChannelDiff<float, 6, int64_t>;
PD_DEFINE_FEATURE(cd_f32x6_i64, T_cd_f32x6_i64)
using T_cd_u8x1_i8 = ChannelDiff<uint8_t, 1, int8_t>;
PD_DEFINE_FEATURE(cd_u8x1_i8, T_cd_u8x1_i8)
