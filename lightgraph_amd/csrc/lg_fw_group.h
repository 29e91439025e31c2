// Pivot groups of the register-tiled symmetric min-plus closures (the pivot tile
// of the blocked Floyd-Warshall and the in-LDS closure of small components).
//
// For the group of pivots k..k+KG-1 let G[r] be row k+r after pivots
// k..k+r-1 (the row the sequential loop uses at pivot k+r); by symmetry it is
// also column k+r.  The KG sequential steps then give exactly
//   D_ij <- min(D_ij, G[0][i] + G[0][j], ..., G[KG-1][i] + G[KG-1][j])
// and G[r][x] = P_r[x] folded with min(., G[q][k+r] + G[q][x]) for q < r, where
// P_r is row k+r before the group: the same sums of the same operands as the
// one-pivot-at-a-time loop, and min is exact, so the result is bit-identical
// to it for any group size (LightGraph.h:512-542 restated on optical depth).
#pragma once

namespace lg {

// Pivots per group: one pair of workgroup barriers per group.
constexpr int kPivotGroup = 8;

// G[r][x] for the group at pivot k, from the published rows P[r] = row k+r
// before the group, by the thread owning column x.  s[q][r] = G[q][k+r] is
// computed in the same order by every thread.
template <int KG, int W>
__device__ __forceinline__ void pivot_group_rows(const float (*P)[W], float (*G)[W], int k, int x) {
    float s[KG][KG];
#pragma unroll
    for (int r = 1; r < KG; ++r) s[0][r] = P[0][k + r];
#pragma unroll
    for (int q = 1; q < KG - 1; ++q)
#pragma unroll
        for (int r = q + 1; r < KG; ++r) {
            float t = P[q][k + r];
#pragma unroll
            for (int p = 0; p < q; ++p) t = fminf(t, s[p][q] + s[p][r]);
            s[q][r] = t;
        }
    float g[KG];
    g[0] = P[0][x];
#pragma unroll
    for (int r = 1; r < KG; ++r) {
        float t = P[r][x];
#pragma unroll
        for (int q = 0; q < r; ++q) t = fminf(t, s[q][r] + g[q]);
        g[r] = t;
    }
#pragma unroll
    for (int r = 0; r < KG; ++r) G[r][x] = g[r];
}

// pivot_group_rows for a 128-column tile computed by one wave (columns lane
// and lane + 64): with 4 x 4 blocks over 32 column groups, the owners of a
// group's 8 rows are exactly one wave, which can then publish its rows and
// derive G without a workgroup barrier in between.  Same operations, same
// order, per column as pivot_group_rows.
template <int KG>
__device__ __forceinline__ void pivot_group_rows_wave128(const float (*P)[128], float (*G)[128], int k, int lane) {
    float s[KG][KG];
#pragma unroll
    for (int r = 1; r < KG; ++r) s[0][r] = P[0][k + r];
#pragma unroll
    for (int q = 1; q < KG - 1; ++q)
#pragma unroll
        for (int r = q + 1; r < KG; ++r) {
            float t = P[q][k + r];
#pragma unroll
            for (int p = 0; p < q; ++p) t = fminf(t, s[p][q] + s[p][r]);
            s[q][r] = t;
        }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int x = lane + 64 * h;
        float g[KG];
        g[0] = P[0][x];
#pragma unroll
        for (int r = 1; r < KG; ++r) {
            float t = P[r][x];
#pragma unroll
            for (int q = 0; q < r; ++q) t = fminf(t, s[q][r] + g[q]);
            g[r] = t;
        }
#pragma unroll
        for (int r = 0; r < KG; ++r) G[r][x] = g[r];
    }
}

// The group's update of the 4 x 4 block at rows 4*rg.., columns 4*cg..: two
// pivots folded per v_min3_f32.
template <int KG, int W>
__device__ __forceinline__ void pivot_group_update(const float (*G)[W], int rg, int cg, float v[4][4]) {
#pragma unroll
    for (int q = 0; q < KG; q += 2) {
        const float4 a0 = *reinterpret_cast<const float4 *>(&G[q][4 * rg]);
        const float4 b0 = *reinterpret_cast<const float4 *>(&G[q][4 * cg]);
        const float4 a1 = *reinterpret_cast<const float4 *>(&G[q + 1][4 * rg]);
        const float4 b1 = *reinterpret_cast<const float4 *>(&G[q + 1][4 * cg]);
        const float ra0[4] = {a0.x, a0.y, a0.z, a0.w}, cb0[4] = {b0.x, b0.y, b0.z, b0.w};
        const float ra1[4] = {a1.x, a1.y, a1.z, a1.w}, cb1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) v[a][b] = fminf(fminf(v[a][b], ra0[a] + cb0[b]), ra1[a] + cb1[b]);
    }
}

}  // namespace lg
