// p16_lanes.h -- Poseidon2-16 with one state element per lane: a 16-lane row of
// a wave holds the state, the external layer's 4x4 blocks and cross-block sums and
// the internal layer's state sum are DPP row operations.  Used where one
// permutation's latency is on a serial chain (Merkle subtrees, the device-side
// transcript).
#pragma once
#include "common.h"

namespace wk {
using namespace kb;

template <int CTRL>
__device__ __forceinline__ F dpp(F x) {
    return F{(uint32_t)__builtin_amdgcn_update_dpp(0, (int)x.v, CTRL, 0xF, 0xF, true)};
}
constexpr int DPP_QXOR1 = 0xB1;    // quad_perm [1,0,3,2]
constexpr int DPP_QXOR2 = 0x4E;    // quad_perm [2,3,0,1]
constexpr int DPP_QNEXT = 0x39;    // quad_perm [1,2,3,0]
constexpr int DPP_ROR4 = 0x124;    // row_ror:4
constexpr int DPP_ROR8 = 0x128;    // row_ror:8
constexpr int DPP_MIRROR = 0x140;  // row_mirror
constexpr int DPP_HMIRROR = 0x141; // row_half_mirror

__device__ __forceinline__ F ext_row16(F x) {
    F s = fadd(x, dpp<DPP_QXOR1>(x));
    s = fadd(s, dpp<DPP_QXOR2>(s));
    const F nx = dpp<DPP_QNEXT>(x);
    const F y = fadd(fadd(s, x), fadd(nx, nx));
    F t = fadd(y, dpp<DPP_ROR4>(y));
    t = fadd(t, dpp<DPP_ROR8>(t));
    return fadd(y, t);
}
__device__ __forceinline__ F row_sum16(F x) {
    F s = fadd(x, dpp<DPP_QXOR1>(x));
    s = fadd(s, dpp<DPP_QXOR2>(s));
    s = fadd(s, dpp<DPP_HMIRROR>(s));
    return fadd(s, dpp<DPP_MIRROR>(s));
}
// A lane's round constants and diagonal entry, loaded once per kernel: a load
// inside the rounds would put a cache round trip on the permutation's serial chain
// (one wave per permutation row, nothing else to hide it).
struct Row16 {
    F ei[4], et[4], d;
    bool first;
};
__device__ __forceinline__ Row16 row16_consts(int i, const P2Consts16 &k) {
    Row16 c;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        c.ei[r] = k.ext_init[r][i];
        c.et[r] = k.ext_term[r][i];
    }
    c.d = F{DDIAG.d16[i]};
    c.first = i == 0;
    return c;
}
// all 64 lanes must call this (DPP reads neighbours); c from row16_consts(lane & 15)
__device__ __forceinline__ F p16_row(F x, const Row16 &c, const P2Consts16 &k) {
    x = ext_row16(x);
#pragma unroll
    for (int r = 0; r < 4; r++) x = ext_row16(fcube(fadd(x, c.ei[r])));
#pragma unroll
    for (int r = 0; r < 20; r++) {
        const F y = fcube(fadd(x, k.internal[r]));
        if (c.first) x = y;
        x = fadd(row_sum16(x), fmul(x, c.d));
    }
#pragma unroll
    for (int r = 0; r < 4; r++) x = ext_row16(fcube(fadd(x, c.et[r])));
    return x;
}


}  // namespace wk
