// LDS tree-merge engine shared by the LDS tier (lds.hip) and the long-row
// group merge (long.hip).
//
// n sorted, duplicate-free lists sit in LDS buffer 0, list k starting at
// loff[0][k] and carrying node id lid[0][k] (its first list position, relative
// to an aligned subtree root).  Level by level, lists whose ids share a parent
// (id >> 1 equal) merge pairwise -- equal columns summed left + right -- and
// the others are carried; merge-path partitioning spreads every level's raw
// elements evenly over the workgroup (count pass, block scan, write pass),
// alternating between the two buffers.  This is the value rule of
// BRMerge-SpGEMM/inc/spgemm.h:639-707 (aligned binary tree over list
// positions, absent children passed through), so sums are bit-identical.
#pragma once
#include "brm_device.h"

namespace brm {

template <int NT, int CAP, int LMAX>
struct MergeSmem {
    double val[2][CAP];
    int col[2][CAP];
    int lid[2][LMAX];
    int loff[2][LMAX + 1];
    int gs[LMAX + 1];
    int sg[NT + 1];
    int si[NT + 1];
    int sj[NT + 1];
    int scan[NT / BRM_WAVE + 1];
    int ngroups;
};

// Merges the n lists held in buffer 0 (node ids sm.lid[0], offsets
// sm.loff[0][0..n]) level by level until one list is left.  Returns its
// length; *cur_out names the buffer holding it.  Uniform across the block.
template <int NT, int CAP, int LMAX>
__device__ int merge_levels(MergeSmem<NT, CAP, LMAX>& sm, int n, int* cur_out) {
    const int t = threadIdx.x;
    int cur = 0;
    while (n > 1) {
        const int nxt = cur ^ 1;
        int G = 0;
        if constexpr (LMAX <= BRM_WAVE) {
            // few lists: one wavefront forms the groups with a ballot
            if (t < BRM_WAVE) {
                const int k = t;
                const bool flag = k < n && (k == 0 || ((sm.lid[cur][k - 1] >> 1) != (sm.lid[cur][k] >> 1)));
                const uint64_t m = __ballot(flag);
                const int ex = __popcll(m & ((1ull << lane_id()) - 1ull));
                if (flag) {
                    sm.gs[ex] = k;
                    sm.lid[nxt][ex] = sm.lid[cur][k] >> 1;
                }
                if (t == 0) {
                    sm.ngroups = __popcll(m);
                    sm.gs[__popcll(m)] = n;
                }
            }
            __syncthreads();
            G = sm.ngroups;
        } else {
            for (int cs = 0; cs < n; cs += NT) {
                const int k = cs + t;
                int flag = 0;
                if (k < n) flag = (k == 0) || ((sm.lid[cur][k - 1] >> 1) != (sm.lid[cur][k] >> 1));
                int tot;
                const int ex = block_excl_scan<NT>(flag, sm.scan, &tot);
                if (flag) {
                    sm.gs[G + ex] = k;
                    sm.lid[nxt][G + ex] = sm.lid[cur][k] >> 1;
                }
                G += tot;
            }
            if (t == 0) sm.gs[G] = n;
            __syncthreads();
        }
        if (G == n) {  // no pair on this level: only the node ids move up
            for (int k = t; k < n; k += NT) sm.lid[cur][k] = sm.lid[nxt][k];
            __syncthreads();
            continue;
        }
        const int* __restrict__ lo_ = sm.loff[cur];
        const int* __restrict__ cc = sm.col[cur];
        const double* __restrict__ cv = sm.val[cur];
        const int Sl = lo_[n];
        // Elements per thread, odd: threads start their merges c elements
        // apart, and an odd stride spreads a wavefront's LDS accesses over
        // all banks (an even one, e.g. 8, lands 16 lanes on one bank).
        const int c = ((Sl + NT - 1) / NT) | 1;

        // split points of threads t (and NT)
        for (int tt = t; tt <= NT; tt += NT) {
            const int d = min(tt * c, Sl);
            // largest g in [0, G] with goff(g) <= d, goff(G) = Sl
            int glo = 0, glen = G + 1;
            while (glen > 1) {
                const int half = glen >> 1;
                const int mid = glo + half;
                const int gm = (mid == G) ? Sl : lo_[sm.gs[mid]];
                if (gm <= d) { glo = mid; glen -= half; } else { glen = half; }
            }
            int i = 0, j = 0;
            if (glo < G) {
                const int k = sm.gs[glo];
                const int L0 = lo_[k], R0 = lo_[k + 1];
                const int nL = R0 - L0;
                const bool pair = (sm.gs[glo + 1] - k) == 2;
                const int nR = pair ? lo_[k + 2] - R0 : 0;
                const int dd = d - L0;
                if (!pair) { i = dd; j = 0; }
                else {
                    merge_split([&](int x) { return cc[L0 + x]; }, nL, [&](int y) { return cc[R0 + y]; }, nR,
                                dd, &i, &j);
                }
            }
            sm.sg[tt] = glo;
            sm.si[tt] = i;
            sm.sj[tt] = j;
        }
        __syncthreads();

        const int g0 = sm.sg[t], i0 = sm.si[t], j0 = sm.sj[t];
        const int ge = sm.sg[t + 1], ie = sm.si[t + 1], je = sm.sj[t + 1];

        // count pass
        int cnt = 0;
        {
            int g = g0, i = i0, j = j0;
            while (g < ge || (g == ge && (i < ie || j < je))) {
                const int k = sm.gs[g];
                const int L0 = lo_[k], R0 = lo_[k + 1];
                const bool pair = (sm.gs[g + 1] - k) == 2;
                const int nL = R0 - L0;
                const int nR = pair ? lo_[k + 2] - R0 : 0;
                const int iend = (g == ge) ? ie : nL;
                const int jend = (g == ge) ? je : nR;
                while (i < iend && j < jend) {
                    const int x = cc[L0 + i], y = cc[R0 + j];
                    i += (x <= y);
                    j += (y <= x);
                    cnt++;
                }
                cnt += (iend - i) + (jend - j);
                g++;
                i = 0;
                j = 0;
            }
        }
        int total;
        const int obase = block_excl_scan_1b<NT>(cnt, sm.scan, &total);

        // write pass
        {
            int* __restrict__ oc = sm.col[nxt];
            double* __restrict__ ov = sm.val[nxt];
            int o = obase;
            int g = g0, i = i0, j = j0;
            while (g < ge || (g == ge && (i < ie || j < je))) {
                const int k = sm.gs[g];
                const int L0 = lo_[k], R0 = lo_[k + 1];
                const bool pair = (sm.gs[g + 1] - k) == 2;
                const int nL = R0 - L0;
                const int nR = pair ? lo_[k + 2] - R0 : 0;
                const int iend = (g == ge) ? ie : nL;
                const int jend = (g == ge) ? je : nR;
                if (i == 0 && j == 0) sm.loff[nxt][g] = o;
                while (i < iend && j < jend) {
                    const int x = cc[L0 + i], y = cc[R0 + j];
                    if (x == y) {
                        oc[o] = x;
                        ov[o] = fadd(cv[L0 + i], cv[R0 + j]);
                        i++;
                        j++;
                    } else if (x < y) {
                        oc[o] = x;
                        ov[o] = cv[L0 + i];
                        i++;
                    } else {
                        oc[o] = y;
                        ov[o] = cv[R0 + j];
                        j++;
                    }
                    o++;
                }
                for (; i < iend; ++i, ++o) { oc[o] = cc[L0 + i]; ov[o] = cv[L0 + i]; }
                for (; j < jend; ++j, ++o) { oc[o] = cc[R0 + j]; ov[o] = cv[R0 + j]; }
                g++;
                i = 0;
                j = 0;
            }
        }
        if (t == 0) sm.loff[nxt][G] = total;
        __syncthreads();
        cur = nxt;
        n = G;
    }

    *cur_out = cur;
    return (n == 1) ? sm.loff[cur][1] : 0;
}

}  // namespace brm
