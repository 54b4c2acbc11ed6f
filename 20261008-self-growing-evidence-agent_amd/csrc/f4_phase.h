// f4_phase.h -- 0/1 bit tiles as e2m1 (fp4) matrix operands, exactly.
//
// gfx950 has no 1-bit matrix instruction; its fp4 MFMAs (v_mfma_scale_f32_*_f8f6f4 with
// e2m1 operands) run at twice the int8 rate. Bit phase p (0..3) of a 32-bit bitmap dword
// takes the transactions at bit p of every nibble: masking with 0x1111..., 0x2222... or
// 0x4444... leaves them as the e2m1 values 0.5, 1.0 or 2.0 (bit 3 alone would be -0.0, so
// phase 3 shifts bit 3 to bit 2 first). The instruction's E8M0 scale operands (2^1, 2^0
// or 2^-1 on both A and B) make every matching pair multiply to exactly 1.0, so a K step
// counts |A_row & B_col| over its transactions in fp32, exactly while sums stay below
// 2^24. A and B use the same (lane, VGPR, nibble) -> transaction map, so the K order
// inside an instruction does not change the sum.
#pragma once
#include <stdint.h>

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

template <int P>
__device__ __forceinline__ int sg_f4_phase(uint32_t x)
{
    if (P == 0) return (int)(x & 0x11111111u);
    if (P == 1) return (int)(x & 0x22222222u);
    if (P == 2) return (int)(x & 0x44444444u);
    return (int)((x >> 1) & 0x44444444u);
}

// E8M0 scale of phase P for each operand, so that a matching pair multiplies to 1.0:
// 0.5 * 0.5 * 2 * 2, 1 * 1, 2 * 2 * 2^-1 * 2^-1.
template <int P>
__device__ __forceinline__ constexpr int sg_f4_scale()
{
    return P == 0 ? 128 : P == 1 ? 127 : 126;
}

// Phase p of four bitmap dwords as an fp4 operand (32 values; the upper four VGPRs of the
// 8-VGPR operand are unused for fp4).
__device__ __forceinline__ v8i sg_f4_mask_frag(const uint4& v, int p)
{
    if (p == 0) return v8i{sg_f4_phase<0>(v.x), sg_f4_phase<0>(v.y), sg_f4_phase<0>(v.z), sg_f4_phase<0>(v.w), 0, 0, 0, 0};
    if (p == 1) return v8i{sg_f4_phase<1>(v.x), sg_f4_phase<1>(v.y), sg_f4_phase<1>(v.z), sg_f4_phase<1>(v.w), 0, 0, 0, 0};
    if (p == 2) return v8i{sg_f4_phase<2>(v.x), sg_f4_phase<2>(v.y), sg_f4_phase<2>(v.z), sg_f4_phase<2>(v.w), 0, 0, 0, 0};
    return v8i{sg_f4_phase<3>(v.x), sg_f4_phase<3>(v.y), sg_f4_phase<3>(v.z), sg_f4_phase<3>(v.w), 0, 0, 0, 0};
}
