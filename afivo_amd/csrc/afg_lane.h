// Lane exchange within a wavefront for the base relaxation's cells (device code
// only): the value of v in lane ^ mask, for the masks 1, 2 and 4 of the
// neighbour cells of a 2^D box held on 2^D consecutive lanes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

namespace afg {

// DPP lane permutations (row_mask = bank_mask = 0xf: every lane written; every
// source lane lies in the lane's own group of 8, so none is out of range)
template <int CTRL>
__device__ inline int dpp_move(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, false);
}

// quad_perm [1,0,3,2]: lane ^ 1; quad_perm [2,3,0,1]: lane ^ 2;
// row_half_mirror (lane 7 - i within 8, i.e. i ^ 7) then quad_perm [3,2,1,0]
// (i ^ 3): lane ^ 4.  A DPP move is a VALU operand modifier, with no LDS round
// trip of ds_bpermute (__shfl_xor).
template <int MASK>
__device__ inline int lane_xor_word(int x)
{
    if constexpr (MASK == 1) return dpp_move<0xB1>(x);
    else if constexpr (MASK == 2) return dpp_move<0x4E>(x);
    else return dpp_move<0x1B>(dpp_move<0x141>(x));
}

template <int MASK, typename T>
__device__ inline T lane_xor_t(T v)
{
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    int w[sizeof(T) / 4];
    std::memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) w[i] = lane_xor_word<MASK>(w[i]);
    std::memcpy(&v, w, sizeof(T));
    return v;
}

// every lane of the wavefront must be active (as for __shfl_xor)
template <typename T>
__device__ inline T lane_xor(T v, int mask)
{
    switch (mask) {
    case 1: return lane_xor_t<1>(v);
    case 2: return lane_xor_t<2>(v);
    case 4: return lane_xor_t<4>(v);
    default: return __shfl_xor(v, mask);
    }
}

}  // namespace afg
