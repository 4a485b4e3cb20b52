// Register merge of one short row (at most G products, one per lane of a
// group of G lanes), shared by the short tier (short.hip) and the fused
// small-product path (small.hip).
//
// Each lane holds one product (column, list position, value).  The products
// are first put in (column, position) order, equal columns adjacent and
// ordered by position:
//  - when the columns fit 32 - log2(G) bits, a bitonic sort of 32-bit keys
//    (column, lane) over the group's lanes -- lanes are in list order, so
//    within one column the lane order is the position order -- with DPP
//    quad permutes for lane distances 1 and 2 and ds_swizzle for 4 to 16;
//    each sorted lane then fetches its value and position from the lane its
//    key names;
//  - otherwise a rank sort by 64-bit (column, position) keys (v_readlane
//    broadcasts, or shuffles inside a half-wavefront), and ds_permute moves
//    every lane's product to its rank.
// The merge tree of BRMerge (BRMerge-SpGEMM/inc/spgemm.h:639-707: lists
// paired (2j, 2j+1) level by level, odd list carried) is then evaluated
// column-wise: at level l two live entries of the same column combine iff
// their positions share the parent node (pos >> (l+1)); the left one keeps
// the sum, the right one dies.  That is exactly the reference's value rule
// -- an aligned binary tree over list positions in which a node with one
// present child passes it through -- so the sums are bit-identical.
#pragma once
#include "brm_device.h"

namespace brm {

__device__ __forceinline__ int permute_i32(int dst_lane, int x) {
    return __builtin_amdgcn_ds_permute(dst_lane << 2, x);
}
__device__ __forceinline__ double permute_f64(int dst_lane, double x) {
    int lo = __builtin_amdgcn_ds_permute(dst_lane << 2, __double2loint(x));
    int hi = __builtin_amdgcn_ds_permute(dst_lane << 2, __double2hiint(x));
    return __hiloint2double(hi, lo);
}

// x of lane (this lane ^ J), within groups of 32 lanes (J < 64).
template <int J>
__device__ __forceinline__ uint32_t xor_lanes(uint32_t x) {
    if constexpr (J == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (J == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else if constexpr (J < 32) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (J << 10));  // bitmask mode, xor J
    } else {
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane_id() ^ J) << 2, (int)x);
    }
}

// One bitonic merge of blocks of K lanes (half-cleaners J = K/2 .. 1).
template <int K, int J>
__device__ __forceinline__ void bitonic_merge(uint32_t& key, int gl) {
    const uint32_t o = xor_lanes<J>(key);
    const bool keep_min = ((gl & J) == 0) == ((gl & K) == 0);
    key = keep_min ? min(key, o) : max(key, o);
    if constexpr (J > 1) bitonic_merge<K, J / 2>(key, gl);
}

// Ascending sort of one key per lane over each group of G lanes.
template <int K, int G>
__device__ __forceinline__ void bitonic_sort(uint32_t& key, int gl) {
    bitonic_merge<K, K / 2>(key, gl);
    if constexpr (K < G) bitonic_sort<K * 2, G>(key, gl);
}

// Whether columns below N fit the 32-bit sort keys of groups of G lanes
// (the all-ones key marks an empty lane).
template <int G>
__host__ __device__ constexpr bool short_keys_fit(int N) {
    return (int64_t)N < ((int64_t)1 << (32 - (G == 64 ? 6 : 5)));
}

// Result for the calling lane: its output entry (col, val) when alive, at
// position idx of the row; nout entries in all.  With NUMERIC false only
// nout is computed (the distinct columns).
struct ShortMerged {
    int col;
    double val;
    bool alive;
    int idx;
    int nout;
};

// Every lane of the wavefront calls this (groups of G lanes; G = 64 or 32).
// gl: lane within the group; gbase: the group's first lane; gmask: its
// lanes; act: the lane holds a product (gl < ip); ip: products of the group's
// row (uniform over the group; for G = 64, over the wavefront); narrow:
// short_keys_fit<G>(N) for the product's N (uniform over the wavefront).
template <bool NUMERIC, int G>
__device__ __forceinline__ ShortMerged short_merge(int col, int p, double val, bool act, int ip, int gl, int gbase,
                                                   uint64_t gmask, bool narrow) {
    const int lane = lane_id();
    constexpr int S = G == 64 ? 6 : 5;
    bool alive = gl < ip;
    int pos;
    if (narrow) {
        uint32_t key = act ? (((uint32_t)col << S) | (uint32_t)gl) : 0xFFFFFFFFu;
        bitonic_sort<2, G>(key, gl);
        const int src = gbase + (int)(key & (G - 1));
        col = (int)(key >> S);
        if (!NUMERIC) {
            const int prv = __shfl(col, (lane + BRM_WAVE - 1) & (BRM_WAVE - 1), BRM_WAVE);
            ShortMerged r{};
            r.nout = __popcll(__ballot(alive && (gl == 0 || prv != col)) & gmask);
            return r;
        }
        pos = __builtin_amdgcn_ds_bpermute(src << 2, p);
        int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(val));
        int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(val));
        val = __hiloint2double(hi, lo);
    } else {
        const uint64_t key = act ? (((uint64_t)(uint32_t)col << 32) | (uint32_t)p) : ~0ull;
        const uint32_t klo_me = (uint32_t)key, khi_me = (uint32_t)(key >> 32);
        int rank = 0;
        bool dup = false;
        if (G == BRM_WAVE) {
            const int n = __builtin_amdgcn_readfirstlane(ip);
            for (int k = 0; k < n; ++k) {
                const uint32_t klo = (uint32_t)__builtin_amdgcn_readlane((int)klo_me, k);
                const uint32_t khi = (uint32_t)__builtin_amdgcn_readlane((int)khi_me, k);
                const uint64_t kk = ((uint64_t)khi << 32) | klo;
                const bool lt = kk < key;
                rank += lt;
                dup |= lt && (khi == khi_me);
            }
        } else {
#pragma unroll 8
            for (int k = 0; k < G; ++k) {
                const uint32_t klo = (uint32_t)__shfl((int)klo_me, gbase + k, BRM_WAVE);
                const uint32_t khi = (uint32_t)__shfl((int)khi_me, gbase + k, BRM_WAVE);
                const uint64_t kk = ((uint64_t)khi << 32) | klo;
                const bool lt = kk < key;
                rank += lt;
                dup |= lt && (khi == khi_me);
            }
        }
        if (!NUMERIC) {
            ShortMerged r{};
            r.nout = __popcll(__ballot(act && !dup) & gmask);
            return r;
        }
        const int dst = act ? gbase + rank : lane;
        col = permute_i32(dst, col);
        pos = permute_i32(dst, p);
        val = permute_f64(dst, val);
    }

    const uint64_t below_me = (1ull << lane) - 1ull;
    const uint64_t above_me = lane == BRM_WAVE - 1 ? 0ull : ~0ull << (lane + 1);
    for (int l = 0; l < 31; ++l) {
        const uint64_t am = __ballot(alive) & gmask;
        const uint64_t above = am & above_me;
        const int nxt = above ? (__ffsll((long long)above) - 1) : lane;
        const int ncol = __shfl(col, nxt, BRM_WAVE);
        const int npos = __shfl(pos, nxt, BRM_WAVE);
        const double nval = __shfl(val, nxt, BRM_WAVE);
        const bool same = alive && above != 0ull && ncol == col;
        if (__ballot(same) == 0ull) break;  // no column left to combine in any group
        const bool comb = same && ((npos >> (l + 1)) == (pos >> (l + 1)));
        const uint64_t cm = __ballot(comb);
        const uint64_t below = am & below_me;
        const int prv = below ? (63 - __clzll((long long)below)) : -1;
        const bool dies = prv >= 0 && ((cm >> prv) & 1ull);
        if (comb) val = fadd(val, nval);
        alive = alive && !dies;
    }
    const uint64_t am = __ballot(alive) & gmask;
    ShortMerged r{};
    r.col = col;
    r.val = val;
    r.alive = alive;
    r.idx = __popcll(am & below_me);
    r.nout = __popcll(am);
    return r;
}

}  // namespace brm
