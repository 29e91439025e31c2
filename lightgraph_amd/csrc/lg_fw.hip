// Kernel (c): blocked Floyd-Warshall in the (min, +) semiring on optical depth
// (the reference's (max, x) on per-channel transmittance, LightGraph.h:512-542).
//
// Schedule for pivot block K (tiles of 128 x 128, pivots in ascending order):
//   1. pivot: close tile (K,K) on one workgroup (1024 threads, registers), four
//      pivots per LDS round trip;
//   2. row panel: (K,J) = min((K,J), P*(K,K) (x) (K,J)) for every J != K, where P*
//      is the closed pivot tile (its zero diagonal keeps the old value); the
//      column panel (J,K) receives the transpose;
//   3. remainder: for I <= J (both != K): (I,J) = min((I,J), (K,I)^T (x) (K,J)),
//      and (J,I) receives the transpose.
// D is symmetric (undirected graph), so phases 2 and 3 read only the row panel
// and compute half of the tiles.  Phases 2 and 3 are one min-plus tile kernel:
// C = min(C, X^T (x) Y) with X, Y row-panel tiles, 8x8 outputs per thread, two
// pivots per step folded with v_min3_f32.  Tiles whose row-panel operand is all
// +INF are skipped (the reference skips the same zero-transmittance entries,
// LightGraph.h:518-520); skipping changes no value.
#include <algorithm>
#include <mutex>
#include <map>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "lg_internal.h"
#include "lg_fw_group.h"

namespace lg {

constexpr int TB = FW_TILE;  // 128
constexpr int KC = 32;       // pivots staged in LDS per step
// Panel launches with fewer half-tile workgroups than this run in column quarters.
constexpr int kQuarterPanelBelow = 256;

// Closure of the symmetric pivot tile, kPivotGroup pivots per round trip through
// LDS (lg_fw_group.h).  Thread t holds the 4 x 4 block (rows 4*(t/32).., cols
// 4*(t%32)..) in registers.
// Pivots of round kb that are vertices (the last round of an n-vertex matrix
// holds n - kb * TB of them; the rest are padding, +INF rows and columns, which
// relax nothing), and the staged pivot chunks that cover them.
__device__ __forceinline__ int fw_round_pivots(int n, int kb) {
    const int m = n - kb * TB;
    return m < TB ? m : TB;
}
__device__ __forceinline__ int fw_round_kend(int n, int kb) {
    return ((fw_round_pivots(n, kb) + KC - 1) / KC) * KC;
}

// m: pivots of the tile to take (the rest are padding)
__device__ __forceinline__ void fw_pivot_tile(float *__restrict__ D, int ld, int kb, int m) {
    constexpr int KG = kPivotGroup;
    static_assert(KG == 8 && TB == 128, "the owners of a pivot group are one wave");
    __shared__ float P[KG][TB];     // rows of the group before it
    __shared__ float G[2][KG][TB];  // the group's pivot rows, alternating buffers
    const int tid = threadIdx.x;
    const int rg = tid >> 5, cg = tid & 31;
    const int wave = tid >> 6, lane = tid & 63;
    float *base = D + (long long)kb * TB * ld + (long long)kb * TB;
    float v[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float4 q = *reinterpret_cast<const float4 *>(base + (long long)(4 * rg + a) * ld + 4 * cg);
        v[a][0] = q.x;
        v[a][1] = q.y;
        v[a][2] = q.z;
        v[a][3] = q.w;
    }
    // Group g's rows k..k+7 are held by wave g (row groups 2g, 2g+1): it
    // publishes them and derives the group's pivot rows itself, so one
    // workgroup barrier per group suffices.  G alternates between two buffers:
    // a wave writes G[g & 1] again only at group g + 2, after every wave has
    // passed group g + 1's barrier and so finished reading it.
    const int groups = (m + KG - 1) / KG;
    for (int g = 0; g < groups; ++g) {
        const int k = KG * g;
        if (wave == g) {
            const int r0 = 4 * (rg % (KG / 4));
#pragma unroll
            for (int a = 0; a < 4; ++a)
                *reinterpret_cast<float4 *>(&P[r0 + a][4 * cg]) = make_float4(v[a][0], v[a][1], v[a][2], v[a][3]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            pivot_group_rows_wave128<KG>(P, G[g & 1], k, lane);
        }
        __syncthreads();
        pivot_group_update<KG, TB>(G[g & 1], rg, cg, v);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
        *reinterpret_cast<float4 *>(base + (long long)(4 * rg + a) * ld + 4 * cg) =
            make_float4(v[a][0], v[a][1], v[a][2], v[a][3]);
}

// One matrix: also clears the tile flags for the round's panel launch.
__global__ void __launch_bounds__(1024) k_fw_pivot(float *__restrict__ D, int ld, int kb, int *__restrict__ flags,
                                                   int nt, int n) {
    for (int t = threadIdx.x; t < nt; t += blockDim.x) flags[t] = 0;
    fw_pivot_tile(D, ld, kb, fw_round_pivots(n, kb));
}

// One matrix of a batch: D (ld x ld, ld = 128 * nt), its tile flags at flags[flag_off].
struct FwJob {
    float *D;
    int ld, nt, flag_off, n;  // n: vertices (rows past n are padding)
};

// Batched pivot closure: block b closes the pivot tile of job active[b].
// It also clears the job's tile flags for the round's panel launch.
__global__ void __launch_bounds__(1024) k_fw_pivot_b(const FwJob *__restrict__ jobs, const int *__restrict__ active,
                                                     int kb, int *__restrict__ flags,
                                                     unsigned long long *__restrict__ stats) {
    const FwJob j = jobs[active[blockIdx.x]];
    if (stats && threadIdx.x == 0)
        stat_add(stats, ST_FW_PIVOTS, (unsigned long long)fw_round_pivots(j.n, kb) * TB * TB);
    for (int t = threadIdx.x; t < j.nt; t += blockDim.x) flags[j.flag_off + t] = 0;
    fw_pivot_tile(j.D, j.ld, kb, fw_round_pivots(j.n, kb));
}

__device__ __forceinline__ void tri_decode(int p, int T, int &a, int &b) {
    // row-major upper triangle of a T x T matrix: row a holds T - a entries
    const double tt = 2.0 * T + 1.0;
    int r = (int)floor((tt - sqrt(tt * tt - 8.0 * (double)p)) * 0.5);
    auto S = [&](int x) { return x * T - (x * (x - 1)) / 2; };
    if (r < 0) r = 0;
    if (r > T - 1) r = T - 1;
    while (r + 1 <= T - 1 && S(r + 1) <= p) ++r;
    while (r > 0 && S(r) > p) --r;
    a = r;
    b = r + (p - S(r));
}

template <bool PANEL>
__device__ __forceinline__ void fw_minplus_tile(float *__restrict__ D, int ld, int kb, int nt, int kend,
                                                int *__restrict__ flags, int tile,
                                                unsigned long long *__restrict__ stats = nullptr) {
    int Ib, Jb;
    if (PANEL) {
        Ib = kb;
        Jb = tile < kb ? tile : tile + 1;
    } else {
        int a, b;
        tri_decode(tile, nt - 1, a, b);
        Ib = a < kb ? a : a + 1;
        Jb = b < kb ? b : b + 1;
        if (!flags[Ib] || !flags[Jb]) return;
    }
    if (!PANEL && stats && threadIdx.x == 0) stat_add(stats, ST_FW_TILES, (unsigned long long)kend * TB * TB);
    __shared__ float Xs[KC][TB];
    __shared__ float Ys[KC][TB];
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    float acc[8][8];
    float *Cb = D + (long long)Ib * TB * ld + (long long)Jb * TB;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int row = (r < 4) ? (ty * 4 + r) : (64 + ty * 4 + r - 4);
        const float4 lo = *reinterpret_cast<const float4 *>(Cb + (long long)row * ld + tx * 4);
        const float4 hi = *reinterpret_cast<const float4 *>(Cb + (long long)row * ld + 64 + tx * 4);
        acc[r][0] = lo.x;
        acc[r][1] = lo.y;
        acc[r][2] = lo.z;
        acc[r][3] = lo.w;
        acc[r][4] = hi.x;
        acc[r][5] = hi.y;
        acc[r][6] = hi.z;
        acc[r][7] = hi.w;
    }
    if (PANEL) {
        // the panel tile is also the Y operand: all +INF leaves it (and its
        // flag, cleared by the caller) unchanged
        int finite = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int t = 0; t < 8; ++t) finite |= (acc[r][t] < INFINITY) ? 1 : 0;
        if (!__syncthreads_or(finite)) return;
        if (stats && threadIdx.x == 0) stat_add(stats, ST_FW_PANEL_TILES, (unsigned long long)kend * TB * TB);
    }
    const float *Xg = D + (long long)kb * TB * ld + (long long)Ib * TB;
    const float *Yg = D + (long long)kb * TB * ld + (long long)Jb * TB;
    for (int kc = 0; kc < kend; kc += KC) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = tid + q * 256;
            const int kk = idx >> 5, c4 = idx & 31;
            *reinterpret_cast<float4 *>(&Xs[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Xg + (long long)(kc + kk) * ld + c4 * 4);
            *reinterpret_cast<float4 *>(&Ys[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Yg + (long long)(kc + kk) * ld + c4 * 4);
        }
        __syncthreads();
#pragma unroll 1  // unrolled, the compiler folds only a third of the pairs into v_min3
        for (int kk = 0; kk < KC; kk += 2) {
            const float4 xa0 = *reinterpret_cast<const float4 *>(&Xs[kk][ty * 4]);
            const float4 xa1 = *reinterpret_cast<const float4 *>(&Xs[kk][64 + ty * 4]);
            const float4 yb0 = *reinterpret_cast<const float4 *>(&Ys[kk][tx * 4]);
            const float4 yb1 = *reinterpret_cast<const float4 *>(&Ys[kk][64 + tx * 4]);
            const float4 xc0 = *reinterpret_cast<const float4 *>(&Xs[kk + 1][ty * 4]);
            const float4 xc1 = *reinterpret_cast<const float4 *>(&Xs[kk + 1][64 + ty * 4]);
            const float4 yd0 = *reinterpret_cast<const float4 *>(&Ys[kk + 1][tx * 4]);
            const float4 yd1 = *reinterpret_cast<const float4 *>(&Ys[kk + 1][64 + tx * 4]);
            const float a[8] = {xa0.x, xa0.y, xa0.z, xa0.w, xa1.x, xa1.y, xa1.z, xa1.w};
            const float b[8] = {yb0.x, yb0.y, yb0.z, yb0.w, yb1.x, yb1.y, yb1.z, yb1.w};
            const float c[8] = {xc0.x, xc0.y, xc0.z, xc0.w, xc1.x, xc1.y, xc1.z, xc1.w};
            const float d[8] = {yd0.x, yd0.y, yd0.z, yd0.w, yd1.x, yd1.y, yd1.z, yd1.w};
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int s = 0; s < 8; ++s) acc[r][s] = fminf(acc[r][s], fminf(a[r] + b[s], c[r] + d[s]));
        }
    }
    // write C
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int row = (r < 4) ? (ty * 4 + r) : (64 + ty * 4 + r - 4);
        *reinterpret_cast<float4 *>(Cb + (long long)row * ld + tx * 4) =
            make_float4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
        *reinterpret_cast<float4 *>(Cb + (long long)row * ld + 64 + tx * 4) =
            make_float4(acc[r][4], acc[r][5], acc[r][6], acc[r][7]);
    }
    if (PANEL || Ib != Jb) {
        // transpose into tile (Jb, Ib)
        float *Tb = D + (long long)Jb * TB * ld + (long long)Ib * TB;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int col = (s < 4) ? (tx * 4 + s) : (64 + tx * 4 + s - 4);
            *reinterpret_cast<float4 *>(Tb + (long long)col * ld + ty * 4) =
                make_float4(acc[0][s], acc[1][s], acc[2][s], acc[3][s]);
            *reinterpret_cast<float4 *>(Tb + (long long)col * ld + 64 + ty * 4) =
                make_float4(acc[4][s], acc[5][s], acc[6][s], acc[7][s]);
        }
    }
    if (PANEL) {
        int finite = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int s = 0; s < 8; ++s) finite |= (acc[r][s] < INFINITY) ? 1 : 0;
        finite = __syncthreads_or(finite);
        if (tid == 0) flags[Jb] = finite;
    }
}

// Panel update of one 128 x 64 column half h (columns 64 h..) of the row-panel
// tile (kb, Jb), 8 x 4 outputs per thread: two workgroups share a panel tile
// when a launch holds too few tiles to fill the GPU.  The panel tile is its own
// Y operand, so a workgroup may only read the columns it writes: a split by
// rows would let one part read rows that another part has already rewritten.
// Pivots fold in the same pairs and order as fw_minplus_tile (bit-identical
// results).  A half that is all +INF stays so and is skipped; a half with a
// finite value sets flags[Jb] to 1 (the caller has cleared the flags).
__device__ __forceinline__ void fw_panel_half(float *__restrict__ D, int ld, int kb, int kend, int *__restrict__ flags,
                                              int tile, int h, unsigned long long *__restrict__ stats) {
    const int Ib = kb, Jb = tile < kb ? tile : tile + 1;
    __shared__ float Xs[KC][TB];
    __shared__ float Ys[KC][64];
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    const int c0 = 64 * h;
    float acc[8][4];
    float *Cb = D + (long long)Ib * TB * ld + (long long)Jb * TB + c0;
    int finite = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int row = (r < 4) ? (ty * 4 + r) : (64 + ty * 4 + r - 4);
        const float4 v = *reinterpret_cast<const float4 *>(Cb + (long long)row * ld + tx * 4);
        acc[r][0] = v.x;
        acc[r][1] = v.y;
        acc[r][2] = v.z;
        acc[r][3] = v.w;
        finite |= (v.x < INFINITY) | (v.y < INFINITY) | (v.z < INFINITY) | (v.w < INFINITY);
    }
    if (!__syncthreads_or(finite)) return;  // all +INF: nothing changes
    if (stats && threadIdx.x == 0) stat_add(stats, ST_FW_PANEL_TILES, (unsigned long long)kend * TB * 64);
    const float *Xg = D + (long long)kb * TB * ld + (long long)Ib * TB;
    const float *Yg = D + (long long)kb * TB * ld + (long long)Jb * TB + c0;
    for (int kc = 0; kc < kend; kc += KC) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = tid + q * 256;
            const int kk = idx >> 5, c4 = idx & 31;
            *reinterpret_cast<float4 *>(&Xs[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Xg + (long long)(kc + kk) * ld + c4 * 4);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int idx = tid + q * 256;
            const int kk = idx >> 4, c4 = idx & 15;
            *reinterpret_cast<float4 *>(&Ys[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Yg + (long long)(kc + kk) * ld + c4 * 4);
        }
        __syncthreads();
#pragma unroll 1  // unrolled, the compiler folds only a third of the pairs into v_min3
        for (int kk = 0; kk < KC; kk += 2) {
            const float4 xa0 = *reinterpret_cast<const float4 *>(&Xs[kk][ty * 4]);
            const float4 xa1 = *reinterpret_cast<const float4 *>(&Xs[kk][64 + ty * 4]);
            const float4 yb = *reinterpret_cast<const float4 *>(&Ys[kk][tx * 4]);
            const float4 xc0 = *reinterpret_cast<const float4 *>(&Xs[kk + 1][ty * 4]);
            const float4 xc1 = *reinterpret_cast<const float4 *>(&Xs[kk + 1][64 + ty * 4]);
            const float4 yd = *reinterpret_cast<const float4 *>(&Ys[kk + 1][tx * 4]);
            const float a[8] = {xa0.x, xa0.y, xa0.z, xa0.w, xa1.x, xa1.y, xa1.z, xa1.w};
            const float b[4] = {yb.x, yb.y, yb.z, yb.w};
            const float c[8] = {xc0.x, xc0.y, xc0.z, xc0.w, xc1.x, xc1.y, xc1.z, xc1.w};
            const float d[4] = {yd.x, yd.y, yd.z, yd.w};
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[r][t] = fminf(acc[r][t], fminf(a[r] + b[t], c[r] + d[t]));
        }
    }
    // every global read of this workgroup happened before the last barrier, so
    // the in-place writes below cannot feed its own loads
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int row = (r < 4) ? (ty * 4 + r) : (64 + ty * 4 + r - 4);
        *reinterpret_cast<float4 *>(Cb + (long long)row * ld + tx * 4) =
            make_float4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
    }
    // transpose into the column panel tile (Jb, kb)
    float *Tb = D + (long long)Jb * TB * ld + (long long)Ib * TB;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const long long col = c0 + tx * 4 + t;
        *reinterpret_cast<float4 *>(Tb + col * ld + ty * 4) = make_float4(acc[0][t], acc[1][t], acc[2][t], acc[3][t]);
        *reinterpret_cast<float4 *>(Tb + col * ld + 64 + ty * 4) =
            make_float4(acc[4][t], acc[5][t], acc[6][t], acc[7][t]);
    }
    finite = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int t = 0; t < 4; ++t) finite |= (acc[r][t] < INFINITY) ? 1 : 0;
    finite = __syncthreads_or(finite);
    if (tid == 0 && finite) flags[Jb] = 1;
}

// fw_panel_half over a 32-column quarter (q = 0..3) of the panel tile, for
// panel launches too small to fill the GPU: the columns of a panel tile are
// independent (column j reads only column j of the tile and the pivot tile),
// and each output folds the pivots in the same pairs and order as
// fw_minplus_tile, so the result is bit-identical.  256 threads, 4 x 4 outputs
// each.
__device__ __forceinline__ void fw_panel_quarter(float *__restrict__ D, int ld, int kb, int kend,
                                                 int *__restrict__ flags, int tile, int q,
                                                 unsigned long long *__restrict__ stats) {
    const int Ib = kb, Jb = tile < kb ? tile : tile + 1;
    __shared__ float Xs[KC][TB];
    __shared__ float Ys[KC][32];
    const int tid = threadIdx.x;
    const int ty = tid >> 3, tx = tid & 7;
    const int c0 = 32 * q;
    float acc[4][4];
    float *Cb = D + (long long)Ib * TB * ld + (long long)Jb * TB + c0;
    int finite = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float4 v = *reinterpret_cast<const float4 *>(Cb + (long long)(ty * 4 + r) * ld + tx * 4);
        acc[r][0] = v.x;
        acc[r][1] = v.y;
        acc[r][2] = v.z;
        acc[r][3] = v.w;
        finite |= (v.x < INFINITY) | (v.y < INFINITY) | (v.z < INFINITY) | (v.w < INFINITY);
    }
    if (!__syncthreads_or(finite)) return;  // all +INF: nothing changes
    if (stats && threadIdx.x == 0) stat_add(stats, ST_FW_PANEL_TILES, (unsigned long long)kend * TB * 32);
    const float *Xg = D + (long long)kb * TB * ld + (long long)Ib * TB;
    const float *Yg = D + (long long)kb * TB * ld + (long long)Jb * TB + c0;
    for (int kc = 0; kc < kend; kc += KC) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = tid + u * 256;
            const int kk = idx >> 5, c4 = idx & 31;
            *reinterpret_cast<float4 *>(&Xs[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Xg + (long long)(kc + kk) * ld + c4 * 4);
        }
        {
            const int kk = tid >> 3, c4 = tid & 7;
            *reinterpret_cast<float4 *>(&Ys[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Yg + (long long)(kc + kk) * ld + c4 * 4);
        }
        __syncthreads();
#pragma unroll 1
        for (int kk = 0; kk < KC; kk += 2) {
            const float4 xa = *reinterpret_cast<const float4 *>(&Xs[kk][ty * 4]);
            const float4 yb = *reinterpret_cast<const float4 *>(&Ys[kk][tx * 4]);
            const float4 xc = *reinterpret_cast<const float4 *>(&Xs[kk + 1][ty * 4]);
            const float4 yd = *reinterpret_cast<const float4 *>(&Ys[kk + 1][tx * 4]);
            const float a[4] = {xa.x, xa.y, xa.z, xa.w};
            const float b[4] = {yb.x, yb.y, yb.z, yb.w};
            const float c[4] = {xc.x, xc.y, xc.z, xc.w};
            const float d[4] = {yd.x, yd.y, yd.z, yd.w};
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[r][t] = fminf(acc[r][t], fminf(a[r] + b[t], c[r] + d[t]));
        }
    }
    // every global read of this workgroup happened before the last barrier, so
    // the in-place writes below cannot feed its own loads
#pragma unroll
    for (int r = 0; r < 4; ++r)
        *reinterpret_cast<float4 *>(Cb + (long long)(ty * 4 + r) * ld + tx * 4) =
            make_float4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
    // transpose into the column panel tile (Jb, kb)
    float *Tb = D + (long long)Jb * TB * ld + (long long)Ib * TB;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const long long col = c0 + tx * 4 + t;
        *reinterpret_cast<float4 *>(Tb + col * ld + ty * 4) = make_float4(acc[0][t], acc[1][t], acc[2][t], acc[3][t]);
    }
    finite = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int t = 0; t < 4; ++t) finite |= (acc[r][t] < INFINITY) ? 1 : 0;
    finite = __syncthreads_or(finite);
    if (tid == 0 && finite) flags[Jb] = 1;
}

// The remainder update for one 64 x 64 quadrant q (rows 64 (q >> 1)..,
// columns 64 (q & 1).. of the output tile), 4 x 4 outputs per thread: four
// workgroups share a tile when a launch holds too few tiles to fill the GPU.
// Its operands are panel tiles, which no remainder workgroup writes.  Each
// output folds the pivots in the same pairs and order as fw_minplus_tile, so
// the result is bit-identical.
__device__ __forceinline__ void fw_minplus_quad(float *__restrict__ D, int ld, int kb, int nt, int kend,
                                                int *__restrict__ flags, int tile, int q,
                                                unsigned long long *__restrict__ stats) {
    int a, b;
    tri_decode(tile, nt - 1, a, b);
    const int Ib = a < kb ? a : a + 1;
    const int Jb = b < kb ? b : b + 1;
    if (!flags[Ib] || !flags[Jb]) return;
    if (stats && threadIdx.x == 0) stat_add(stats, ST_FW_TILES, (unsigned long long)kend * 64 * 64);
    __shared__ float Xs[KC][64];
    __shared__ float Ys[KC][64];
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    const int r0 = 64 * (q >> 1), c0 = 64 * (q & 1);
    float acc[4][4];
    float *Cb = D + (long long)Ib * TB * ld + (long long)Jb * TB;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float4 v = *reinterpret_cast<const float4 *>(Cb + (long long)(r0 + ty * 4 + r) * ld + c0 + tx * 4);
        acc[r][0] = v.x;
        acc[r][1] = v.y;
        acc[r][2] = v.z;
        acc[r][3] = v.w;
    }
    const float *Xg = D + (long long)kb * TB * ld + (long long)Ib * TB + r0;
    const float *Yg = D + (long long)kb * TB * ld + (long long)Jb * TB + c0;

    for (int kc = 0; kc < kend; kc += KC) {
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = tid + h * 256;
            const int kk = idx >> 4, c4 = idx & 15;
            *reinterpret_cast<float4 *>(&Xs[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Xg + (long long)(kc + kk) * ld + c4 * 4);
            *reinterpret_cast<float4 *>(&Ys[kk][c4 * 4]) =
                *reinterpret_cast<const float4 *>(Yg + (long long)(kc + kk) * ld + c4 * 4);
        }
        __syncthreads();
#pragma unroll 1  // unrolled, the compiler folds only a third of the pairs into v_min3
        for (int kk = 0; kk < KC; kk += 2) {
            const float4 xa = *reinterpret_cast<const float4 *>(&Xs[kk][ty * 4]);
            const float4 yb = *reinterpret_cast<const float4 *>(&Ys[kk][tx * 4]);
            const float4 xc = *reinterpret_cast<const float4 *>(&Xs[kk + 1][ty * 4]);
            const float4 yd = *reinterpret_cast<const float4 *>(&Ys[kk + 1][tx * 4]);
            const float a[4] = {xa.x, xa.y, xa.z, xa.w};
            const float b[4] = {yb.x, yb.y, yb.z, yb.w};
            const float c[4] = {xc.x, xc.y, xc.z, xc.w};
            const float d[4] = {yd.x, yd.y, yd.z, yd.w};
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[r][t] = fminf(acc[r][t], fminf(a[r] + b[t], c[r] + d[t]));
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
        *reinterpret_cast<float4 *>(Cb + (long long)(r0 + ty * 4 + r) * ld + c0 + tx * 4) =
            make_float4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
    if (Ib != Jb) {
        float *Tb = D + (long long)Jb * TB * ld + (long long)Ib * TB;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            *reinterpret_cast<float4 *>(Tb + (long long)(c0 + tx * 4 + t) * ld + r0 + ty * 4) =
                make_float4(acc[0][t], acc[1][t], acc[2][t], acc[3][t]);
    }
}

// Workgroups per 128 x 128 tile: one, or split (SPLIT 2) into two panel column
// halves or four remainder quadrants.  (Remainder halves measured slower than
// quadrants at every launch size.)
template <bool PANEL, int SPLIT>
constexpr int fw_parts() {
    return SPLIT == 1 ? 1 : (PANEL ? (SPLIT == 4 ? 4 : 2) : 4);
}

template <bool PANEL, int SPLIT>
__global__ void __launch_bounds__(256) k_fw_minplus(float *__restrict__ D, int ld, int kb, int nt,
                                                     int *__restrict__ flags, int n) {
    constexpr int parts = fw_parts<PANEL, SPLIT>();
    const int kend = fw_round_kend(n, kb);
    if constexpr (SPLIT == 1)
        fw_minplus_tile<PANEL>(D, ld, kb, nt, kend, flags, (int)blockIdx.x);
    else if constexpr (PANEL)
        fw_panel_half(D, ld, kb, kend, flags, (int)blockIdx.x / parts, (int)blockIdx.x % parts, nullptr);
    else
        fw_minplus_quad(D, ld, kb, nt, kend, flags, (int)blockIdx.x / parts, (int)blockIdx.x % parts, nullptr);
}

// Batched panel / remainder launch: prefix[a] is the first tile of active job
// a; fw_parts workgroups per tile.
template <bool PANEL, int SPLIT>
__global__ void __launch_bounds__(256) k_fw_minplus_b(const FwJob *__restrict__ jobs, const int *__restrict__ active,
                                                       const int *__restrict__ prefix, int n_active, int kb,
                                                       int *__restrict__ flags, unsigned long long *__restrict__ stats) {
    constexpr int parts = fw_parts<PANEL, SPLIT>();
    const int b = (int)blockIdx.x / parts;
    int lo = 0, hi = n_active - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const FwJob j = jobs[active[lo]];
    const int kend = fw_round_kend(j.n, kb);
    if constexpr (SPLIT == 1)
        fw_minplus_tile<PANEL>(j.D, j.ld, kb, j.nt, kend, flags + j.flag_off, b - prefix[lo], stats);
    else if constexpr (PANEL && SPLIT == 4)
        fw_panel_quarter(j.D, j.ld, kb, kend, flags + j.flag_off, b - prefix[lo], (int)blockIdx.x % parts, stats);
    else if constexpr (PANEL)
        fw_panel_half(j.D, j.ld, kb, kend, flags + j.flag_off, b - prefix[lo], (int)blockIdx.x % parts, stats);
    else
        fw_minplus_quad(j.D, j.ld, kb, j.nt, kend, flags + j.flag_off, b - prefix[lo], (int)blockIdx.x % parts,
                        stats);
}

// ------------------------------------------------------------------ tuning ----
// The closures's two choices (lg_fw_tuning): which launches split tiles
// into quadrants, and whether the rounds run as a captured hipGraph.
struct FwTuningChoice {
    int split_below = 0;
    int use_graph = 0;
};
static int32_t fw_tuning_active(lg_ctx *ctx, FwTuningChoice *out);

int32_t fw_solve_impl(lg_ctx *ctx, int32_t n, float *d_D, int32_t ld) {
    LG_CHECK_ARG(ld % TB == 0 && ld >= n && n >= 0, "fw: ld=%d must be a multiple of %d and >= n=%d", ld, TB, n);
    if (n <= 1) return LG_OK;
    const int nt = ld / TB;
    void *s = nullptr;
    LG_TRY(scratch(ctx, (size_t)nt * sizeof(int), &s));
    int *flags = (int *)s;
    const int T = nt - 1;
    const long long P = (long long)T * (T + 1) / 2;
    LG_CHECK_ARG(P < (1LL << 31), "fw: matrix too large");
    FwTuningChoice tc;
    LG_TRY(fw_tuning_active(ctx, &tc));
    for (int kb = 0; kb < nt; ++kb) {
        hipEvent_t e1, e2, e3;
        ktimer_begin(ctx, "k_fw_pivot", &e1);
        k_fw_pivot<<<1, 1024, 0, ctx->stream>>>(d_D, ld, kb, flags, nt, n);
        LG_LAUNCHED("k_fw_pivot");
        ktimer_end(ctx, "k_fw_pivot", e1);
        if (nt == 1) break;
        ktimer_begin(ctx, "k_fw_panel", &e2);
        if (nt - 1 < tc.split_below)
            k_fw_minplus<true, 2><<<2 * (nt - 1), 256, 0, ctx->stream>>>(d_D, ld, kb, nt, flags, n);
        else
            k_fw_minplus<true, 1><<<nt - 1, 256, 0, ctx->stream>>>(d_D, ld, kb, nt, flags, n);
        LG_LAUNCHED("k_fw_panel");
        ktimer_end(ctx, "k_fw_panel", e2);
        ktimer_begin(ctx, "k_fw_minplus", &e3);
        if (P < tc.split_below && 4 * P < (1LL << 31))
            k_fw_minplus<false, 2><<<(unsigned)(4 * P), 256, 0, ctx->stream>>>(d_D, ld, kb, nt, flags, n);
        else
            k_fw_minplus<false, 1><<<(unsigned)P, 256, 0, ctx->stream>>>(d_D, ld, kb, nt, flags, n);
        LG_LAUNCHED("k_fw_minplus");
        ktimer_end(ctx, "k_fw_minplus", e3);
    }
    return LG_OK;
}

template <typename F>
static int32_t fw_graph_capture(lg_ctx *ctx, const std::vector<uint64_t> &key, F &&rounds);
static int32_t fw_batched_run(lg_ctx *ctx, const std::vector<float *> &Ds, const std::vector<int> &ns,
                              unsigned long long *stats, const FwTuningChoice &tc);

int32_t fw_solve_batched_impl(lg_ctx *ctx, const std::vector<float *> &Ds, const std::vector<int> &ns,
                              unsigned long long *stats) {
    FwTuningChoice tc;
    LG_TRY(fw_tuning_active(ctx, &tc));
    return fw_batched_run(ctx, Ds, ns, stats, tc);
}

// Uneven batch: round kb runs the pivot tiles of every matrix with more than kb
// tile rows in one launch, then all their panels, then all their remainders.
static int32_t fw_batched_run(lg_ctx *ctx, const std::vector<float *> &Ds, const std::vector<int> &ns,
                              unsigned long long *stats, const FwTuningChoice &tc) {
    const int count = (int)Ds.size();
    std::vector<FwJob> jobs;
    int max_nt = 0, flag_total = 0;
    for (int g = 0; g < count; ++g) {
        if (ns[g] <= 1) continue;
        const int ld = fw_ld(ns[g]);
        jobs.push_back(FwJob{Ds[g], ld, ld / TB, flag_total, ns[g]});
        flag_total += ld / TB;
        max_nt = std::max(max_nt, ld / TB);
    }
    if (jobs.empty()) return LG_OK;
    const int J = (int)jobs.size();
    // per round: active job ids, panel prefix, remainder prefix
    std::vector<int> table;
    std::vector<int> round_off(max_nt), round_n(max_nt), panel_blocks(max_nt), rem_blocks(max_nt);
    for (int kb = 0; kb < max_nt; ++kb) {
        round_off[kb] = (int)table.size();
        std::vector<int> act;
        for (int j = 0; j < J; ++j)
            if (jobs[j].nt > kb) act.push_back(j);
        round_n[kb] = (int)act.size();
        table.insert(table.end(), act.begin(), act.end());
        int pp = 0, rp = 0;
        std::vector<int> ppre, rpre;
        for (int j : act) {
            ppre.push_back(pp);
            rpre.push_back(rp);
            const long long T = jobs[j].nt - 1;
            pp += (int)T;
            rp += (int)(T * (T + 1) / 2);
        }
        panel_blocks[kb] = pp;
        rem_blocks[kb] = rp;
        table.insert(table.end(), ppre.begin(), ppre.end());
        table.insert(table.end(), rpre.begin(), rpre.end());
    }
    const size_t jobs_bytes = sizeof(FwJob) * J;
    const size_t table_bytes = sizeof(int) * table.size();
    const size_t flag_bytes = sizeof(int) * (size_t)flag_total;
    void *s = nullptr;
    LG_TRY(scratch(ctx, jobs_bytes + table_bytes + flag_bytes + 1024, &s));
    const size_t table_at = ((jobs_bytes + 255) / 256) * 256;
    FwJob *d_jobs = (FwJob *)s;
    int *d_table = (int *)((char *)s + table_at);
    int *d_flags = d_table + table.size() + 64;
    // the job list and the round tables in one upload
    std::vector<uint8_t> up(table_at + table_bytes, 0);
    memcpy(up.data(), jobs.data(), jobs_bytes);
    memcpy(up.data() + table_at, table.data(), table_bytes);
    LG_HIP(hipMemcpyAsync(s, up.data(), up.size(), hipMemcpyHostToDevice, ctx->stream));
    auto rounds = [&](bool timers) -> int32_t {
        for (int kb = 0; kb < max_nt; ++kb) {
            const int *act = d_table + round_off[kb];
            const int n = round_n[kb];
            hipEvent_t e1 = nullptr, e2 = nullptr, e3 = nullptr;
            if (timers) ktimer_begin(ctx, "k_fw_pivot", &e1);
            k_fw_pivot_b<<<n, 1024, 0, ctx->stream>>>(d_jobs, act, kb, d_flags, stats);
            LG_LAUNCHED("k_fw_pivot_b");
            if (timers) ktimer_end(ctx, "k_fw_pivot", e1);
            if (panel_blocks[kb] == 0) continue;
            if (timers) ktimer_begin(ctx, "k_fw_panel", &e2);
            const bool split_panel = panel_blocks[kb] < tc.split_below && panel_blocks[kb] < (1 << 29);
            if (split_panel && 2 * panel_blocks[kb] < kQuarterPanelBelow)
                // too few halves to fill the GPU: column quarters
                k_fw_minplus_b<true, 4><<<4 * panel_blocks[kb], 256, 0, ctx->stream>>>(d_jobs, act, act + n, n, kb,
                                                                                     d_flags, stats);
            else if (split_panel)
                k_fw_minplus_b<true, 2><<<2 * panel_blocks[kb], 256, 0, ctx->stream>>>(d_jobs, act, act + n, n, kb,
                                                                                     d_flags, stats);
            else
                k_fw_minplus_b<true, 1><<<panel_blocks[kb], 256, 0, ctx->stream>>>(d_jobs, act, act + n, n, kb,
                                                                                 d_flags, stats);
            LG_LAUNCHED("k_fw_panel_b");
            if (timers) ktimer_end(ctx, "k_fw_panel", e2);
            if (timers) ktimer_begin(ctx, "k_fw_minplus", &e3);
            if (rem_blocks[kb] < tc.split_below && rem_blocks[kb] < (1 << 29))
                k_fw_minplus_b<false, 2><<<4 * rem_blocks[kb], 256, 0, ctx->stream>>>(d_jobs, act, act + 2 * n, n,
                                                                                     kb, d_flags, stats);
            else
                k_fw_minplus_b<false, 1><<<rem_blocks[kb], 256, 0, ctx->stream>>>(d_jobs, act, act + 2 * n, n, kb,
                                                                                 d_flags, stats);
            LG_LAUNCHED("k_fw_minplus_b");
            if (timers) ktimer_end(ctx, "k_fw_minplus", e3);
        }
        return LG_OK;
    };
    if (!tc.use_graph || ctx->timing) return rounds(ctx->timing);
    // graph schedule: the launches depend on the device tables' address, the
    // round shapes, the split choice and the counter pointer; the tables'
    // contents are read when the graph runs, so one capture serves every batch
    // of the same shape
    std::vector<uint64_t> key = {(uint64_t)(uintptr_t)s, (uint64_t)(uintptr_t)stats, (uint64_t)tc.split_below,
                                 (uint64_t)max_nt, (uint64_t)table_at};
    for (int kb = 0; kb < max_nt; ++kb) {
        key.push_back((uint64_t)round_off[kb]);
        key.push_back(((uint64_t)round_n[kb] << 32) | (uint32_t)panel_blocks[kb]);
        key.push_back((uint64_t)rem_blocks[kb]);
    }
    if (!ctx->fw_graph || key != ctx->fw_graph_key) LG_TRY(fw_graph_capture(ctx, key, rounds));
    LG_HIP(hipGraphLaunch(ctx->fw_graph, ctx->stream));
    return LG_OK;
}

// Replaces the context's captured schedule by a capture of `rounds`.
template <typename F>
static int32_t fw_graph_capture(lg_ctx *ctx, const std::vector<uint64_t> &key, F &&rounds) {
    if (ctx->fw_graph) {
        LG_HIP(hipStreamSynchronize(ctx->stream));  // the old graph may still run
        (void)hipGraphExecDestroy(ctx->fw_graph);
        ctx->fw_graph = nullptr;
        ctx->fw_graph_key.clear();
    }
    LG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int32_t rc = rounds(false);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != LG_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != LG_OK) return rc;
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return LG_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(ei));
        return LG_ERR_HIP;
    }
    ctx->fw_graph = exec;
    ctx->fw_graph_key = key;
    return LG_OK;
}

// Defaults without a measurement: split launches of fewer than two tiles per CU.
constexpr int kFwSplitBelowDefault = 256;

__global__ void k_tune_fill(float *__restrict__ D, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    D[i] = 1.0f + (float)(h & 1023u) * (1.0f / 1024.0f);
}

__global__ void k_tune_flags(int *__restrict__ f, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = 1;
}

static int env_int(const char *name, int dflt) {
    const char *v = getenv(name);
    return (v && *v) ? atoi(v) : dflt;
}

// Measures, on synthetic finite matrices: (1) one remainder launch of T tiles
// with one and with four workgroups per tile, T = the remainder tiles of round 0
// of an nt-tile matrix, giving split_below = the first T at which splitting no
// longer wins; (2) the closure of 24 matrices of 2..7 tile rows launched on the
// stream and as a captured graph, giving use_graph when a capture pays for
// itself in one use.
static int32_t fw_autotune_run(lg_ctx *ctx, lg_fw_tuning *t) {
    *t = lg_fw_tuning{};
    t->split_below = kFwSplitBelowDefault;
    static const int nts[] = {2, 3, 4, 6, 8, 12, 16, 23, 32, 46, 64, 91};
    const int npts = (int)(sizeof(nts) / sizeof(nts[0]));
    const int ld = 91 * TB;
    float *D = nullptr;
    int *meta = nullptr;
    if (dalloc(ctx, (size_t)ld * ld, &D) != LG_OK) return LG_OK;  // no room: keep the defaults
    if (dalloc(ctx, 256, &meta) != LG_OK) {
        dfree(ctx, D);
        return LG_OK;
    }
    hipStream_t st = ctx->stream;
    const long long nel = (long long)ld * ld;
    k_tune_fill<<<(unsigned)((nel + 255) / 256), 256, 0, st>>>(D, nel);
    k_tune_flags<<<1, 64, 0, st>>>(meta + 64, 64);
    hipEvent_t a = nullptr, b = nullptr;
    int32_t rc = LG_OK;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) rc = LG_ERR_HIP;
    // meta: [0] FwJob, [32] active = {0}, [33] prefix = {0}, [64..] flags
    auto time_launch = [&](auto launch) -> double {
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {
            (void)hipEventRecord(a, st);
            launch();
            (void)hipEventRecord(b, st);
            (void)hipEventSynchronize(b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            if (rep > 0) best = std::min(best, ms);  // rep 0 warms up
        }
        return (double)best * 1e3;
    };
    for (int k = 0; k < npts && rc == LG_OK; ++k) {
        const int nt = nts[k];
        const int T = (nt - 1) * nt / 2;
        FwJob job{D, ld, nt, 0, nt * TB};
        int ap[2] = {0, 0};
        if (hipMemcpyAsync(meta, &job, sizeof(job), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(meta + 32, ap, sizeof(ap), hipMemcpyHostToDevice, st) != hipSuccess) {
            rc = LG_ERR_HIP;
            break;
        }
        const FwJob *dj = (const FwJob *)meta;
        const int *act = meta + 32;
        t->tiles[k] = T;
        t->whole_us[k] = time_launch([&] {
            k_fw_minplus_b<false, 1><<<T, 256, 0, st>>>(dj, act, act + 1, 1, 0, meta + 64, nullptr);
        });
        t->split_us[k] = time_launch([&] {
            k_fw_minplus_b<false, 2><<<4 * T, 256, 0, st>>>(dj, act, act + 1, 1, 0, meta + 64, nullptr);
        });
        t->points = k + 1;
    }
    if (rc == LG_OK && hipGetLastError() != hipSuccess) rc = LG_ERR_HIP;
    if (rc == LG_OK) {
        int sb = 1 << 30;  // splitting won at every measured size
        for (int k = 0; k < t->points; ++k)
            if (t->whole_us[k] <= t->split_us[k]) {
                sb = t->tiles[k];
                break;
            }
        t->split_below = sb;
    }
    // (2) stream against graph, 24 matrices of 2..7 tile rows carved from D
    if (rc == LG_OK) {
        std::vector<float *> Ds;
        std::vector<int> ns;
        size_t off = 0;
        for (int g = 0; g < 24; ++g) {
            const int n = TB * (2 + g % 6);
            Ds.push_back(D + off);
            ns.push_back(n);
            off += (size_t)n * n;
        }
        FwTuningChoice tc;
        tc.split_below = t->split_below;
        float best = 1e30f;
        for (int rep = 0; rep < 4 && rc == LG_OK; ++rep) {
            (void)hipEventRecord(a, st);
            rc = fw_batched_run(ctx, Ds, ns, nullptr, tc);
            (void)hipEventRecord(b, st);
            (void)hipEventSynchronize(b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            if (rep > 0) best = std::min(best, ms);
        }
        t->stream_ms = best;
        tc.use_graph = 1;
        const bool timing = ctx->timing;
        ctx->timing = false;
        if (rc == LG_OK) {
            (void)hipStreamSynchronize(st);
            const auto c0 = std::chrono::steady_clock::now();
            rc = fw_batched_run(ctx, Ds, ns, nullptr, tc);  // captures
            (void)hipStreamSynchronize(st);
            const double first = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
            best = 1e30f;
            for (int rep = 0; rep < 3 && rc == LG_OK; ++rep) {
                (void)hipEventRecord(a, st);
                rc = fw_batched_run(ctx, Ds, ns, nullptr, tc);  // replays
                (void)hipEventRecord(b, st);
                (void)hipEventSynchronize(b);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, a, b);
                best = std::min(best, ms);
            }
            t->graph_ms = best;
            t->capture_ms = std::max(0.0, first - (double)best);
        }
        ctx->timing = timing;
        if (ctx->fw_graph) {
            (void)hipStreamSynchronize(st);
            (void)hipGraphExecDestroy(ctx->fw_graph);
            ctx->fw_graph = nullptr;
            ctx->fw_graph_key.clear();
        }
        t->use_graph = (rc == LG_OK && t->graph_ms + t->capture_ms < t->stream_ms) ? 1 : 0;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    dfree(ctx, meta);
    dfree(ctx, D);
    if (rc != LG_OK) return rc;
    t->tuned = 1;
    return LG_OK;
}

static std::mutex g_fw_tune_mu;
static std::map<int, lg_fw_tuning> g_fw_tuned;  // per device, measured once per process

int32_t fw_tuning_init(lg_ctx *ctx) {
    if (ctx->fw_tune_set) return LG_OK;
    lg_fw_tuning t{};
    t.split_below = kFwSplitBelowDefault;
    if (env_int("LG_FW_AUTOTUNE", 1) != 0) {
        std::lock_guard<std::mutex> lk(g_fw_tune_mu);
        auto it = g_fw_tuned.find(ctx->device);
        if (it != g_fw_tuned.end()) {
            t = it->second;
        } else {
            lg_fw_tuning m{};
            if (fw_autotune_run(ctx, &m) == LG_OK && m.tuned) {
                t = m;
                g_fw_tuned[ctx->device] = m;
            }
        }
    }
    ctx->fw_tune = t;
    ctx->fw_tune_set = true;
    return LG_OK;
}

int32_t fw_autotune(lg_ctx *ctx, lg_fw_tuning *out) {
    std::lock_guard<std::mutex> lk(g_fw_tune_mu);
    LG_TRY(fw_autotune_run(ctx, out));
    if (out->tuned) g_fw_tuned[ctx->device] = *out;
    return LG_OK;
}

static int32_t fw_tuning_active(lg_ctx *ctx, FwTuningChoice *out) {
    LG_TRY(fw_tuning_init(ctx));
    out->split_below = env_int("LG_FW_SPLIT_BELOW", ctx->fw_tune.split_below);
    out->use_graph = env_int("LG_FW_GRAPH", ctx->fw_tune.use_graph) != 0 ? 1 : 0;
    return LG_OK;
}

// ------------------------------------------------------------ predecessors ----
__global__ void k_adj_degree(const int32_t *__restrict__ edges, int ne, int32_t *__restrict__ deg) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    atomicAdd(&deg[edges[2 * e]], 1);
    atomicAdd(&deg[edges[2 * e + 1]], 1);
}

__global__ void k_adj_fill(const int32_t *__restrict__ edges, const float *__restrict__ w, int ne,
                           const int32_t *__restrict__ off, int32_t *__restrict__ cursor, int32_t *__restrict__ nb,
                           float *__restrict__ nw) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int i = edges[2 * e], j = edges[2 * e + 1];
    const float we = w[e];
    int p = off[i] + atomicAdd(&cursor[i], 1);
    nb[p] = j;
    nw[p] = we;
    p = off[j] + atomicAdd(&cursor[j], 1);
    nb[p] = i;
    nw[p] = we;
}

__global__ void k_adj_sort(int n, const int32_t *__restrict__ off, int32_t *__restrict__ nb, float *__restrict__ nw) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int s = off[v], e = off[v + 1];
    for (int a = s + 1; a < e; ++a) {
        const int kn = nb[a];
        const float kw = nw[a];
        int b = a - 1;
        while (b >= s && nb[b] > kn) {
            nb[b + 1] = nb[b];
            nw[b + 1] = nw[b];
            --b;
        }
        nb[b + 1] = kn;
        nw[b + 1] = kw;
    }
}

// pred[i][j] = smallest neighbour k of j minimising D[i][k] + w(k,j).
__global__ void k_pred(int n, int ld, const float *__restrict__ D, const int32_t *__restrict__ off,
                       const int32_t *__restrict__ nb, const float *__restrict__ nw, int32_t *__restrict__ pred) {
    const int i = blockIdx.x;
    const float *row = D + (long long)i * ld;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        int best_k = -1;
        if (j == i) {
            best_k = i;
        } else if (row[j] < INFINITY) {
            float best = INFINITY;
            for (int p = off[j]; p < off[j + 1]; ++p) {
                const int k = nb[p];
                const float v = row[k] + nw[p];
                if (v < best) {
                    best = v;
                    best_k = k;
                }
            }
        }
        pred[(long long)i * ld + j] = best_k;
    }
}

__global__ void k_scan_small(const int32_t *__restrict__ in, int n, int32_t *__restrict__ out) {
    // exclusive scan, single thread block of 1024 (n up to a few 1e5)
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int s = t * per, e = min(n, s + per);
    int32_t acc = 0;
    for (int k = s; k < e; ++k) acc += in[k];
    part[t] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = t ? part[t - 1] : 0;
    for (int k = s; k < e; ++k) {
        out[k] = run;
        run += in[k];
    }
    if (t == 1023) out[n] = part[1023];
}

int32_t fw_pred_impl(lg_ctx *ctx, int32_t n, const float *d_D, int32_t ld, int32_t *d_pred, const int32_t *d_edges,
                     const float *d_w, int32_t ne) {
    int32_t *deg = nullptr, *off = nullptr, *cur = nullptr, *nb = nullptr;
    float *nw = nullptr;
    int32_t rc = dalloc(ctx, (size_t)n + 1, &deg);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)n + 1, &off);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)n + 1, &cur);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)2 * ne + 1, &nb);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)2 * ne + 1, &nw);
    if (rc == LG_OK) {
        (void)hipMemsetAsync(deg, 0, ((size_t)n + 1) * 4, ctx->stream);
        (void)hipMemsetAsync(cur, 0, ((size_t)n + 1) * 4, ctx->stream);
        if (ne > 0) k_adj_degree<<<(ne + 255) / 256, 256, 0, ctx->stream>>>(d_edges, ne, deg);
        k_scan_small<<<1, 1024, 0, ctx->stream>>>(deg, n, off);
        if (ne > 0) k_adj_fill<<<(ne + 255) / 256, 256, 0, ctx->stream>>>(d_edges, d_w, ne, off, cur, nb, nw);
        k_adj_sort<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, off, nb, nw);
        k_pred<<<n, 256, 0, ctx->stream>>>(n, ld, d_D, off, nb, nw, d_pred);
        if (hipGetLastError() != hipSuccess) {
            set_error("predecessor kernels failed to launch");
            rc = LG_ERR_HIP;
        }
    }
    dfree(ctx, deg);
    dfree(ctx, off);
    dfree(ctx, cur);
    dfree(ctx, nb);
    dfree(ctx, nw);
    return rc;
}

}  // namespace lg

using namespace lg;

extern "C" int32_t lg_fw_solve(lg_ctx *ctx, int32_t n, float *d_D, int32_t ld, int32_t *d_pred,
                               const int32_t *d_edges, const float *d_w, int32_t n_edges) {
    LG_CHECK_ARG(ctx && d_D, "lg_fw_solve: NULL argument");
    LG_CHECK_ARG(ld == fw_ld(n), "lg_fw_solve: ld must be lg_fw_ld(n)=%d, got %d", fw_ld(n), ld);
    LG_TRY(fw_solve_impl(ctx, n, d_D, ld));
    if (d_pred) {
        LG_CHECK_ARG(n_edges >= 0 && (n_edges == 0 || (d_edges && d_w)), "lg_fw_solve: predecessors need edges and weights");
        LG_TRY(fw_pred_impl(ctx, n, d_D, ld, d_pred, d_edges, d_w, n_edges));
    }
    return LG_OK;
}
