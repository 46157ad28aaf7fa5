// Jacobi3D FP64 7-point sweep (Cube topology, constant edge in the halo).
//
// Cell semantics follow JacobiCellFixedHood::update
// (src/testbed/performancetests/main.cpp:1194-1205):
//     T' = (T[z-1] + T[y-1] + T[x-1] + T + T[x+1] + T[y+1] + T[z+1]) * (1.0/7.0)
// summed left to right in exactly that order so results are bit-identical
// to a serial CPU sweep.  The reference launches one thread per (x,y)
// column walking all of z (cudasimulator.h:160-236); here a work-group is a
// 64-wide x row segment times `by` rows, each thread marches a bounded z
// chunk keeping the (z-1, z, z+1) column values in registers, so every
// element of the source plane is fetched from HBM once per sweep and the
// x/y neighbours come from the same cache lines.
#include "lgdhip_internal.h"

#include <algorithm>

namespace {

constexpr int kWaveX = 64;  // lanes along x: one 512-byte row segment

template <int BY>
__global__ __launch_bounds__(kWaveX *BY) void jacobi3d_f64_zmarch(
    const double *__restrict__ src, double *__restrict__ dst, int nx, int ny, int nz, int64_t sx,
    int64_t sxy, int64_t origin, int zchunk) {
    const int x = blockIdx.x * kWaveX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int z0 = blockIdx.z * zchunk;
    if (x >= nx || y >= ny) return;
    int z1 = z0 + zchunk;
    if (z1 > nz) z1 = nz;
    int64_t i = origin + (int64_t)z0 * sxy + (int64_t)y * sx + x;
    double below = src[i - sxy];
    double center = src[i];
    for (int z = z0; z < z1; ++z) {
        const double above = src[i + sxy];
        double s = below;   // ( 0,  0, -1)
        s += src[i - sx];   // ( 0, -1,  0)
        s += src[i - 1];    // (-1,  0,  0)
        s += center;        // ( 0,  0,  0)
        s += src[i + 1];    // ( 1,  0,  0)
        s += src[i + sx];   // ( 0,  1,  0)
        s += above;         // ( 0,  0,  1)
        dst[i] = s * (1.0 / 7.0);
        below = center;
        center = above;
        i += sxy;
    }
}

// Same sweep, but the z march advances U planes per iteration and issues all
// 5U loads of a batch before the arithmetic, so each thread keeps U HBM
// fetches (the z+1 planes) in flight instead of one: the single-plane march
// above is latency bound (one outstanding plane per thread).  The sum order
// per cell is unchanged.
template <int BY, int U>
__global__ __launch_bounds__(kWaveX *BY) void jacobi3d_f64_zbatch(
    const double *__restrict__ src, double *__restrict__ dst, int nx, int ny, int nz, int64_t sx,
    int64_t sxy, int64_t origin, int zchunk) {
    const int x = blockIdx.x * kWaveX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int z0 = blockIdx.z * zchunk;
    if (x >= nx || y >= ny) return;
    int z1 = z0 + zchunk;
    if (z1 > nz) z1 = nz;
    int64_t i = origin + (int64_t)z0 * sxy + (int64_t)y * sx + x;
    // q[0] = plane z-1, q[1] = plane z, q[2..U+1] = planes z+1..z+U
    double q[U + 2];
    q[0] = src[i - sxy];
    q[1] = src[i];
    int z = z0;
    for (; z + U <= z1; z += U) {
        double ym[U], xm[U], xp[U], yp[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int64_t j = i + (int64_t)k * sxy;
            q[k + 2] = src[j + sxy];
            ym[k] = src[j - sx];
            xm[k] = src[j - 1];
            xp[k] = src[j + 1];
            yp[k] = src[j + sx];
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            double s = q[k];  // ( 0,  0, -1)
            s += ym[k];       // ( 0, -1,  0)
            s += xm[k];       // (-1,  0,  0)
            s += q[k + 1];    // ( 0,  0,  0)
            s += xp[k];       // ( 1,  0,  0)
            s += yp[k];       // ( 0,  1,  0)
            s += q[k + 2];    // ( 0,  0,  1)
            __builtin_nontemporal_store(s * (1.0 / 7.0), dst + i + (int64_t)k * sxy);
        }
        q[0] = q[U];
        q[1] = q[U + 1];
        i += (int64_t)U * sxy;
    }
    double below = q[0], center = q[1];
    for (; z < z1; ++z) {
        const double above = src[i + sxy];
        double s = below;
        s += src[i - sx];
        s += src[i - 1];
        s += center;
        s += src[i + 1];
        s += src[i + sx];
        s += above;
        dst[i] = s * (1.0 / 7.0);
        below = center;
        center = above;
        i += sxy;
    }
}

// The four-plane batched march with two cells per thread: every load is 16
// bytes (the z planes and y rows aligned, the x neighbours as the pairs
// starting one cell to the left and one to the right, unaligned), so a wave
// covers a 128-cell (1 KiB) row segment with half the load instructions of
// the one-cell march, and the results leave as non-temporal 16-byte stores.
// Needs an even origin index, row and plane pitch and 16-byte aligned
// buffers (a whole grid from the library's allocator); the second cell of
// the last pair on an odd row is computed from the slack cell past the
// halo and not stored.  Per-cell sum order as above.
typedef double dpair __attribute__((ext_vector_type(2)));
typedef double dpair_u __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ dpair ld_pair(const double *p) { return *reinterpret_cast<const dpair *>(p); }
__device__ __forceinline__ dpair ld_pair_u(const double *p) { return *reinterpret_cast<const dpair_u *>(p); }

template <int BY, int U>
__global__ __launch_bounds__(kWaveX *BY) void jacobi3d_f64_zbatch_pair(
    const double *__restrict__ src, double *__restrict__ dst, int nx, int ny, int nz, int64_t sx,
    int64_t sxy, int64_t origin, int zchunk) {
    const int x = 2 * (int)(blockIdx.x * kWaveX + threadIdx.x);
    const int y = blockIdx.y * BY + threadIdx.y;
    const int z0 = blockIdx.z * zchunk;
    if (x >= nx || y >= ny) return;
    const bool both = x + 1 < nx;
    int z1 = z0 + zchunk;
    if (z1 > nz) z1 = nz;
    int64_t i = origin + (int64_t)z0 * sxy + (int64_t)y * sx + x;
    auto put = [&](int64_t at, dpair v) {
        if (both)
            __builtin_nontemporal_store(v, reinterpret_cast<dpair *>(dst + at));
        else
            __builtin_nontemporal_store(v.x, dst + at);
    };
    dpair q[U + 2];
    q[0] = ld_pair(src + i - sxy);
    q[1] = ld_pair(src + i);
    int z = z0;
    for (; z + U <= z1; z += U) {
        dpair ym[U], xm[U], xp[U], yp[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int64_t j = i + (int64_t)k * sxy;
            q[k + 2] = ld_pair(src + j + sxy);
            ym[k] = ld_pair(src + j - sx);
            xm[k] = ld_pair_u(src + j - 1);
            xp[k] = ld_pair_u(src + j + 1);
            yp[k] = ld_pair(src + j + sx);
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            dpair s = q[k];  // ( 0,  0, -1)
            s += ym[k];      // ( 0, -1,  0)
            s += xm[k];      // (-1,  0,  0)
            s += q[k + 1];   // ( 0,  0,  0)
            s += xp[k];      // ( 1,  0,  0)
            s += yp[k];      // ( 0,  1,  0)
            s += q[k + 2];   // ( 0,  0,  1)
            put(i + (int64_t)k * sxy, s * (1.0 / 7.0));
        }
        q[0] = q[U];
        q[1] = q[U + 1];
        i += (int64_t)U * sxy;
    }
    dpair below = q[0], center = q[1];
    for (; z < z1; ++z) {
        const dpair above = ld_pair(src + i + sxy);
        dpair s = below;
        s += ld_pair(src + i - sx);
        s += ld_pair_u(src + i - 1);
        s += center;
        s += ld_pair_u(src + i + 1);
        s += ld_pair(src + i + sx);
        s += above;
        put(i, s * (1.0 / 7.0));
        below = center;
        center = above;
        i += sxy;
    }
}

// value of lane-1 / lane+1 of a double (DPP wave_shr:1 / wave_shl:1 on both
// halves; lanes at the wave's ends read 0 and only feed halo cells)
__device__ __forceinline__ double lane_prev(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_next(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// `inside ? value : edge` as a select: the empty asm makes `value` (and
// the LDS reads feeding it) unconditional, which keeps the compiler from
// turning the select into a branch around the sum -- a branch into which it
// sinks each level's LDS reads and then waits for every read on its own.
// Used by the two-rows kernel (4% faster at 512^3 with its LDS reads hoisted
// to the start of the step); the one-row kernel measures faster with the
// compiler's branch.
__device__ __forceinline__ double keep_select(double value, bool inside, double edge) {
    asm volatile("" : "+v"(value));
    return inside ? value : edge;
}

// Coordinates of the temporal-blocking kernels.  A tile reaches K cells
// past the box; on a wrapped axis (Torus topologies) those cells are the
// periodic images of interior cells and are computed like them, on the
// others they hold the edge value.  Level-0 loads read the image on a
// wrapped axis and clamp into the one-cell halo shell on the others.
__device__ __forceinline__ bool in_axis(int c, int n, int wraps) { return wraps || (unsigned)c < (unsigned)n; }
__device__ __forceinline__ int level0_coord(int c, int n, int wraps) {
    return wraps ? ((c % n) + n) % n : min(max(c, -1), n);
}
// planes run from K before a z chunk to K + PF past it: one period's
// correction reaches them when nz >= 8 (the host's condition for wrapped z)
__device__ __forceinline__ int level0_plane(int p, int n, int wraps) {
    if (wraps) return p < 0 ? p + n : (p >= n ? p - n : p);
    return min(max(p, -1), n);
}

// K sweeps per pass over HBM (temporal blocking).  A work-group is a 64 x TY
// tile of (x, y) columns whose outer K lanes/rows are an overlap ring
// recomputed by the neighbouring tiles, so it writes (64-2K) x (TY-2K)
// columns.  It marches its z chunk as a pipeline: when level-0 plane p
// arrives, level l (l = 1..K) is computed at plane p-l from level l-1 at
// planes p-l-1, p-l, p-l+1 (a 3-plane register window per level); x
// neighbours come from the adjacent lanes by DPP, y neighbours from the
// level's plane p-l published in LDS in the previous step (double
// buffered: one barrier per plane).  Cells outside the box hold the edge
// value at every level, exactly as the halo does between single sweeps, and
// every cell sums in the single sweep's order, so K passes of this kernel
// equal K single sweeps bit for bit.  Level-0 planes are prefetched PF = 3
// planes ahead.
//
// Tiles are numbered x fastest, then y, then z chunk; work-groups map to
// tiles through the XCD swizzle (each XCD a contiguous tile range), so the
// overlap rings of neighbouring tiles are re-read from the same L2.
template <int K, int TY, int PF>
__global__ __launch_bounds__(64 * TY) void jacobi3d_f64_tb(const double *__restrict__ src, double *__restrict__ dst,
                                                           int nx, int ny, int nz, int64_t sx, int64_t sxy,
                                                           int64_t origin, int zchunk, unsigned gx, unsigned gy,
                                                           unsigned ntiles, unsigned per, int swizzle, int wrap,
                                                           int64_t edge_at, int zbase, int zend) {
    constexpr int OX = 64 - 2 * K, OY = TY - 2 * K;
    static_assert(PF == 3, "level-0 ring of 6: a window of 3 planes plus 3 in flight");
    __shared__ double lds[2][K][TY][64];
    const unsigned t = lgdhip::xcd_tile(blockIdx.x, per, swizzle != 0);
    if (t >= ntiles) return;  // whole work-group, before any barrier
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = (int)(t % gx) * OX - K + tx;
    const int y = (int)((t / gx) % gy) * OY - K + ty;
    const int z0 = zbase + (int)(t / (gx * gy)) * zchunk;  // output planes [zbase, zend) in chunks
    const int z1 = min(z0 + zchunk, zend);
    // a cell counts as inside when every coordinate is in range or its axis wraps
    const bool inxy = in_axis(x, nx, wrap & 1) && in_axis(y, ny, wrap & 2);
    const bool writer = (unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny && tx >= K && tx < 64 - K &&
                        ty >= K && ty < TY - K;
    const double edge = src[edge_at];  // a halo cell of an unwrapped axis: the edge value
    const int64_t col = origin + (int64_t)y * sx + x;
    const int yd = ty > 0 ? ty - 1 : 0;
    const int yu = ty < TY - 1 ? ty + 1 : TY - 1;

    // level-0 loads are unconditional: coordinates outside the box map to
    // their periodic image on a wrapped axis and are clamped into the
    // one-cell halo shell (which holds the edge value) on the others, so no
    // select or branch guards them and the compiler's wait counts let PF
    // planes stay in flight
    const int64_t colh = origin + (int64_t)level0_coord(y, ny, wrap & 2) * sx + level0_coord(x, nx, wrap & 1);
    auto load0 = [&](int p) -> double { return src[colh + (int64_t)level0_plane(p, nz, wrap & 4) * sxy]; };

    // lv: level-0 planes, a ring of 6 (plane p in slot (p - pbeg) mod 6:
    // the window p-2 .. p plus the loads of p+1 .. p+3 in flight); w[l]
    // (l >= 1): level l's last three planes, a ring of 3.  The slot roles
    // follow the step index j mod 6, a compile-time constant once the march
    // is unrolled by 6, so no value is ever moved between registers.
    double lv[6];
    double w[K][3];
#pragma unroll
    for (int l = 0; l < K; ++l) w[l][0] = w[l][1] = w[l][2] = edge;
#pragma unroll
    for (int l = 0; l < K; ++l) lds[0][l][ty][tx] = edge;
    const int pbeg = z0 - K;  // level-0 planes pbeg .. z1 + K - 1: 2K ramp steps, then z1 - z0 steps
    lv[4] = lv[5] = edge;
#pragma unroll
    for (int d = 0; d < PF; ++d) lv[d] = load0(pbeg + d);
    __syncthreads();

    // step j at level-0 plane p = pbeg + j; JM = j mod 6; JR = j during the
    // ramp, 2K after.  Level l's plane p - l is read by a written plane only
    // when j >= 2l, and level l's LDS copy only when j >= 2l + 1: the ramp
    // skips the rest (no later step reads those slots, so their stale
    // contents never matter).
    auto step = [&](auto jm_c, auto jr_c, int p) {
        constexpr int JM = decltype(jm_c)::value;
        constexpr int JR = decltype(jr_c)::value;
        constexpr int NEW = JM % 3, OLD = (JM + 1) % 3, MID = (JM + 2) % 3;  // roles after this step's insert
        constexpr int B = JM & 1;                                             // LDS buffer holding step j-1's planes
        lv[(JM + PF) % 6] = load0(p + PF);
#pragma unroll
        for (int l = 1; l <= K; ++l) {
            if (2 * l > JR) continue;
            // level l at plane q = p - l from level l-1 at planes q-1, q, q+1
            const int q = p - l;
            const double a = l == 1 ? lv[(JM + 4) % 6] : w[l - 1][OLD];
            const double c = l == 1 ? lv[(JM + 5) % 6] : w[l - 1][MID];
            const double b = l == 1 ? lv[JM] : w[l - 1][NEW];
            const double xm = lane_prev(c);
            const double xp = lane_next(c);
            const double ym = lds[B][l - 1][yd][tx];
            const double yp = lds[B][l - 1][yu][tx];
            double s = a;  // ( 0,  0, -1)
            s += ym;       // ( 0, -1,  0)
            s += xm;       // (-1,  0,  0)
            s += c;        // ( 0,  0,  0)
            s += xp;       // ( 1,  0,  0)
            s += yp;       // ( 0,  1,  0)
            s += b;        // ( 0,  0,  1)
            // a plain select: the compiler's branch around the sum measures
            // faster here than the forced select of the two-rows kernel
            const double v = (inxy && in_axis(q, nz, wrap & 4)) ? s * (1.0 / 7.0) : edge;
            if (l < K) {
                w[l][NEW] = v;
            } else if (writer) {  // after the ramp q = p - K runs over z0 .. z1 - 1
                __builtin_nontemporal_store(v, dst + col + (int64_t)q * sxy);
            }
        }
        // publish level l at plane p - l for the next step's y neighbours
        if (JR >= 1) lds[B ^ 1][0][ty][tx] = lv[JM];
#pragma unroll
        for (int l = 1; l < K; ++l)
            if (2 * l + 1 <= JR) lds[B ^ 1][l][ty][tx] = w[l][NEW];
        __syncthreads();
    };
    using lgdhip::ic;
    lgdhip::static_for<0, 2 * K>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        step(ic<j % 6>{}, jc, pbeg + j);
    });
    constexpr int S = (2 * K) % 6;
    const int m = z1 - z0;  // steps after the ramp
    int p = z0 + K;
    int i = 0;
    for (; i + 6 <= m; i += 6, p += 6) {
        lgdhip::static_for<0, 6>([&](auto dc) {
            constexpr int d = decltype(dc)::value;
            step(ic<(S + d) % 6>{}, ic<2 * K>{}, p + d);
        });
    }
    lgdhip::static_for<0, 5>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        if (i + d < m) step(ic<(S + d) % 6>{}, ic<2 * K>{}, p + d);
    });
}

// Same K-sweep pipeline with two adjacent rows per thread: a 64 x TY
// work-group holds a 64 x 2TY tile (written region (64-2K) x (2TY-2K)).  A
// thread's rows 2ty and 2ty+1 are each other's y neighbours in registers;
// only the tile-row pairs' outer neighbours come through LDS (the first row
// of thread ty+1 and the last row of thread ty-1), which halves the LDS
// reads per cell, and the taller tile recomputes a smaller overlap ring.
// Cells, sum order, ramp and rings as in jacobi3d_f64_tb: bit-exact with K
// single sweeps.
template <int K, int TY>
__global__ __launch_bounds__(64 * TY) void jacobi3d_f64_tb2(const double *__restrict__ src, double *__restrict__ dst,
                                                            int nx, int ny, int nz, int64_t sx, int64_t sxy,
                                                            int64_t origin, int zchunk, unsigned gx, unsigned gy,
                                                            unsigned ntiles, unsigned per, int swizzle, int wrap,
                                                            int64_t edge_at, int zbase, int zend) {
    constexpr int PF = 3;
    constexpr int CY = 2 * TY;  // tile rows
    constexpr int OX = 64 - 2 * K, OY = CY - 2 * K;
    static_assert(OY > 0, "tile taller than its overlap ring");
    // [buffer][level][thread row][0: first row, 1: last row of the pair][lane]
    __shared__ double lds[2][K][TY][2][64];
    const unsigned t = lgdhip::xcd_tile(blockIdx.x, per, swizzle != 0);
    if (t >= ntiles) return;  // whole work-group, before any barrier
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = (int)(t % gx) * OX - K + tx;
    const int y0 = (int)((t / gx) % gy) * OY - K + 2 * ty;  // rows y0, y0 + 1
    const int z0 = zbase + (int)(t / (gx * gy)) * zchunk;
    const int z1 = min(z0 + zchunk, zend);
    const bool xin = in_axis(x, nx, wrap & 1);
    const bool inxy[2] = {xin && in_axis(y0, ny, wrap & 2), xin && in_axis(y0 + 1, ny, wrap & 2)};
    const bool xw = (unsigned)x < (unsigned)nx && tx >= K && tx < 64 - K;
    const bool writer[2] = {(unsigned)y0 < (unsigned)ny && xw && 2 * ty >= K && 2 * ty < CY - K,
                            (unsigned)(y0 + 1) < (unsigned)ny && xw && 2 * ty + 1 >= K && 2 * ty + 1 < CY - K};
    const double edge = src[edge_at];  // a halo cell of an unwrapped axis: the edge value
    const int64_t col[2] = {origin + (int64_t)y0 * sx + x, origin + (int64_t)(y0 + 1) * sx + x};
    const int xc = level0_coord(x, nx, wrap & 1);
    const int64_t colh[2] = {origin + (int64_t)level0_coord(y0, ny, wrap & 2) * sx + xc,
                             origin + (int64_t)level0_coord(y0 + 1, ny, wrap & 2) * sx + xc};
    // the ring rows (tile rows 0 and CY-1) read their own pair instead of a
    // missing neighbour: their values are never written
    const int tyd = ty > 0 ? ty - 1 : 0;
    const int tyu = ty < TY - 1 ? ty + 1 : TY - 1;
    auto load0 = [&](int r, int p) -> double { return src[colh[r] + (int64_t)level0_plane(p, nz, wrap & 4) * sxy]; };

    double lv[2][6];
    double w[K][2][3];
#pragma unroll
    for (int l = 0; l < K; ++l)
#pragma unroll
        for (int r = 0; r < 2; ++r) w[l][r][0] = w[l][r][1] = w[l][r][2] = edge;
#pragma unroll
    for (int l = 0; l < K; ++l) lds[0][l][ty][0][tx] = lds[0][l][ty][1][tx] = edge;
    const int pbeg = z0 - K;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        lv[r][4] = lv[r][5] = edge;
#pragma unroll
        for (int d = 0; d < PF; ++d) lv[r][d] = load0(r, pbeg + d);
    }
    __syncthreads();

    auto step = [&](auto jm_c, auto jr_c, int p) {
        constexpr int JM = decltype(jm_c)::value;
        constexpr int JR = decltype(jr_c)::value;
        constexpr int NEW = JM % 3, OLD = (JM + 1) % 3, MID = (JM + 2) % 3;
        constexpr int B = JM & 1;
#pragma unroll
        for (int r = 0; r < 2; ++r) lv[r][(JM + PF) % 6] = load0(r, p + PF);
        // the pair's outer y neighbours of every level (previous step's LDS
        // planes), read before the level chain
        double ymv[K], ypv[K];
#pragma unroll
        for (int l = 1; l <= K; ++l) {
            if (2 * l > JR) continue;
            ymv[l - 1] = lds[B][l - 1][tyd][1][tx];
            ypv[l - 1] = lds[B][l - 1][tyu][0][tx];
        }
#pragma unroll
        for (int l = 1; l <= K; ++l) {
            if (2 * l > JR) continue;
            const int q = p - l;
            const bool zin = in_axis(q, nz, wrap & 4);
            double a[2], c[2], b[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                a[r] = l == 1 ? lv[r][(JM + 4) % 6] : w[l - 1][r][OLD];
                c[r] = l == 1 ? lv[r][(JM + 5) % 6] : w[l - 1][r][MID];
                b[r] = l == 1 ? lv[r][JM] : w[l - 1][r][NEW];
            }
            const double ym[2] = {ymv[l - 1], c[0]};
            const double yp[2] = {c[1], ypv[l - 1]};
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                double s = a[r];          // ( 0,  0, -1)
                s += ym[r];               // ( 0, -1,  0)
                s += lane_prev(c[r]);     // (-1,  0,  0)
                s += c[r];                // ( 0,  0,  0)
                s += lane_next(c[r]);     // ( 1,  0,  0)
                s += yp[r];               // ( 0,  1,  0)
                s += b[r];                // ( 0,  0,  1)
                const double v = keep_select(s * (1.0 / 7.0), inxy[r] && zin, edge);
                if (l < K) {
                    w[l][r][NEW] = v;
                } else if (writer[r]) {
                    __builtin_nontemporal_store(v, dst + col[r] + (int64_t)q * sxy);
                }
            }
        }
        if (JR >= 1) {
            lds[B ^ 1][0][ty][0][tx] = lv[0][JM];
            lds[B ^ 1][0][ty][1][tx] = lv[1][JM];
        }
#pragma unroll
        for (int l = 1; l < K; ++l)
            if (2 * l + 1 <= JR) {
                lds[B ^ 1][l][ty][0][tx] = w[l][0][NEW];
                lds[B ^ 1][l][ty][1][tx] = w[l][1][NEW];
            }
        __syncthreads();
    };
    using lgdhip::ic;
    lgdhip::static_for<0, 2 * K>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        step(ic<j % 6>{}, jc, pbeg + j);
    });
    constexpr int S = (2 * K) % 6;
    const int m = z1 - z0;
    int p = z0 + K;
    int i = 0;
    for (; i + 6 <= m; i += 6, p += 6) {
        lgdhip::static_for<0, 6>([&](auto dc) {
            constexpr int d = decltype(dc)::value;
            step(ic<(S + d) % 6>{}, ic<2 * K>{}, p + d);
        });
    }
    lgdhip::static_for<0, 5>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        if (i + d < m) step(ic<(S + d) % 6>{}, ic<2 * K>{}, p + d);
    });
}

// variant ids (lgdhip_tile::variant): 0 default, 1 single-plane march,
// 2 four-plane batches, 3 eight-plane batches; temporal blocking (K sweeps
// per launch, tile 64 x TY): 4 K=2 TY=16, 5 K=3 TY=16, 6 K=4 TY=16,
// 7 K=2 TY=8, 8 K=3 TY=8; two rows per thread (tile 64 x 2TY, work-group
// 64 x TY): 9 K=4 TY=8, 10 K=4 TY=16; the same pipelines for one sweep per
// launch (a plane fetched once per tile, y neighbours through LDS, x by
// DPP): 11 one row per thread, 12 two rows; 13 four-plane batches of two
// cells per thread (16-byte loads; whole grids only, else variant 2).  The
// default blocks 4 sweeps per
// launch with variant 6 or 10, whichever the launch cost model rates cheaper
// for the grid (10 from 96^3, 6 at 64^3, the faster ones there on MI355X);
// step counts not divisible by K finish with single sweeps of
// kSingleSweepVariant.
constexpr int kJacobiDefaultVariant = 6;
constexpr int kSingleSweepVariant = 13;
constexpr int kPairVariant = 13;

struct TbShape {
    int k, ty, rows;  // sweeps per launch, work-group rows, tile rows per work-group row
};
TbShape tb_shape(int variant) {
    switch (variant) {
    case 4: return {2, 16, 1};
    case 5: return {3, 16, 1};
    case 6: return {4, 16, 1};
    case 7: return {2, 8, 1};
    case 8: return {3, 8, 1};
    case 9: return {4, 8, 2};
    case 10: return {4, 16, 2};
    case 11: return {1, 16, 1};
    case 12: return {1, 16, 2};
    default: return {0, 0, 1};
    }
}
constexpr int kJacobiVariants = 13;
constexpr int kJacobiTallVariant = 10;  // the default's choice for grids that fill several rounds

struct JacobiLaunch {
    dim3 grid, block;
    int zchunk;
    int by;
    int variant;
    unsigned gx, gy, ntiles, per;  // temporal blocking: tile grid and XCD swizzle
    int swizzle;
    int zbase = 0, zend = 0;       // temporal blocking: output planes [zbase, zend) (zend 0: all)
};

// Temporal blocking: tiles of (64-2K) x (TY*rows-2K) output columns; a z
// chunk of C planes costs its work-group C + 2K pipeline steps.  Two
// one-row work-groups share a CU (64 KB of LDS each at K = 4, TY = 16), a
// two-rows work-group (128 KB, 102 VGPRs) has it to itself.  A step costs
// (MI355X, 7-point FP64, 128^3 and 512^3 chunk sweeps): one row 0.78 us
// with the CU to itself and 1.17 us with two work-groups on it; two rows
// 1.0 us alone and 1.17 us once the grid is larger than the chip (memory
// traffic of every CU).  A launch takes ceil(work-groups / slots) such
// rounds.  Returns that cost for the best chunk (any C up to 64) and sets
// *zchunk: at 128^3 two rows in chunks of 10 planes (234 work-groups, one
// round of 18 steps, against one row's 480 work-groups in chunks of 13 at
// 21 shared steps); at 512^3 two rows in chunks of 64.
double tb_cost(const lgdhip_grid &g, const TbShape &tb, int *zchunk) {
    const unsigned gx = lgdhip::ceil_div(g.nx, 64 - 2 * tb.k);
    const unsigned gy = lgdhip::ceil_div(g.ny, tb.ty * tb.rows - 2 * tb.k);
    const uint64_t slots = (uint64_t)lgdhip::kCUs * (tb.rows == 2 ? 1 : 2);
    double best = 1e300;
    for (int64_t n = lgdhip::ceil_div(g.nz, 64); n <= g.nz; ++n) {
        const int c = (int)lgdhip::ceil_div(g.nz, n);
        if (lgdhip::ceil_div(g.nz, c) != n) continue;  // same chunk as a smaller n
        const uint64_t groups = (uint64_t)gx * gy * n;
        const bool full = groups > (uint64_t)lgdhip::kCUs;
        const double step = tb.rows == 2 ? (full ? 1.17 : 1.0) : (full ? 1.17 : 0.78);
        const double cost = (double)((groups + slots - 1) / slots) * (c + 2 * tb.k) * step;
        if (cost < best) {
            best = cost;
            *zchunk = c;
        }
    }
    return best;
}

JacobiLaunch plan(const lgdhip_grid &g, const lgdhip_tile *tile) {
    JacobiLaunch L;
    L.by = 4;
    L.zchunk = 16;
    L.variant = kJacobiDefaultVariant;
    if (tile) {
        if (tile->by == 1 || tile->by == 2 || tile->by == 4 || tile->by == 8) L.by = tile->by;
        if (tile->bz > 0) L.zchunk = tile->bz;
        if (tile->variant >= 1 && tile->variant <= kJacobiVariants) L.variant = tile->variant;
    }
    if (L.variant == kJacobiDefaultVariant && !(tile && tile->variant != 0)) {
        // the default: 4 sweeps per launch, one or two rows per thread by the
        // cost model (two rows win once the grid fills several rounds)
        int c6 = 0, c10 = 0;
        if (tb_cost(g, tb_shape(kJacobiTallVariant), &c10) < tb_cost(g, tb_shape(kJacobiDefaultVariant), &c6))
            L.variant = kJacobiTallVariant;
    }
    const TbShape tb = tb_shape(L.variant);
    if (tb.k > 0) {
        const unsigned gx = lgdhip::ceil_div(g.nx, 64 - 2 * tb.k);
        const unsigned gy = lgdhip::ceil_div(g.ny, tb.ty * tb.rows - 2 * tb.k);
        if (!tile || tile->bz <= 0) tb_cost(g, tb, &L.zchunk);
        L.by = tb.ty;
        L.block = dim3(64, tb.ty, 1);
        L.gx = gx;
        L.gy = gy;
        L.ntiles = gx * gy * lgdhip::ceil_div(g.nz, L.zchunk);
        L.swizzle = lgdhip::xcd_swizzle(tile, true);
        L.per = lgdhip::xcd_per(L.ntiles);
        L.grid = dim3(L.swizzle ? lgdhip::xcd_groups(L.ntiles) : L.ntiles, 1, 1);
        return L;
    }
    // the pair march needs even pitches and origin (a whole grid); views
    // and other layouts take the one-cell four-plane batches
    if (L.variant == kPairVariant &&
        ((g.flags & LGDHIP_GRID_VIEW) || (lgdhip::origin_index(g) & 1) || (g.sx & 1) || ((g.sx * g.sy) & 1)))
        L.variant = 2;
    // the pair march: 8-plane chunks (2% faster than 16 at 256^3 and 512^3)
    if (L.variant == kPairVariant && !(tile && tile->bz > 0)) L.zchunk = 8;
    // aim for several thousand work-groups so all 256 CUs stay busy even on
    // small cubes: shrink the z chunk until the grid is large enough
    const unsigned gx = lgdhip::ceil_div(g.nx, L.variant == kPairVariant ? 2 * kWaveX : kWaveX);
    const unsigned gy = lgdhip::ceil_div(g.ny, L.by);
    if (!tile || tile->bz <= 0) {
        while (L.zchunk > 4 && (uint64_t)gx * gy * lgdhip::ceil_div(g.nz, L.zchunk) < 2048) L.zchunk /= 2;
    }
    L.block = dim3(kWaveX, L.by, 1);
    L.grid = dim3(gx, gy, lgdhip::ceil_div(g.nz, L.zchunk));
    return L;
}

template <int BY>
void launch_batched(const double *src, double *dst, const lgdhip_grid &g, const JacobiLaunch &L, hipStream_t s) {
    const int64_t sxy = g.sx * g.sy;
    const int64_t origin = lgdhip::origin_index(g);
    if (L.variant == kPairVariant && !(((uintptr_t)src | (uintptr_t)dst) & 15)) {
        jacobi3d_f64_zbatch_pair<BY, 4><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx,
                                                                   sxy, origin, L.zchunk);
        return;
    }
    if (L.variant == kPairVariant) {  // unaligned buffers: one cell per thread
        const dim3 grid(lgdhip::ceil_div(g.nx, kWaveX), L.grid.y, L.grid.z);
        jacobi3d_f64_zbatch<BY, 4><<<grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                           origin, L.zchunk);
        return;
    }
    if (L.variant == 3)
        jacobi3d_f64_zbatch<BY, 8><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                             origin, L.zchunk);
    else
        jacobi3d_f64_zbatch<BY, 4><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                             origin, L.zchunk);
}

// one launch of the temporal-blocked kernel: K sweeps src -> dst; wrap bits
// 1, 2, 4: x, y, z taken periodic; edge_at: a cell of src holding the edge
// value (a halo cell of an unwrapped axis; any cell when all three wrap)
int launch_tb(const double *src, double *dst, const lgdhip_grid &g, const JacobiLaunch &L, hipStream_t s,
              int wrap = 0, int64_t edge_at = -1) {
    const int64_t sxy = g.sx * g.sy;
    const int64_t origin = lgdhip::origin_index(g);
    if (edge_at < 0) edge_at = origin - 1;
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    const int zend = L.zend > 0 ? L.zend : nz;
#define LGD_TB_ARGS                                                                                                \
    src, dst, nx, ny, nz, g.sx, sxy, origin, L.zchunk, L.gx, L.gy, L.ntiles, L.per, L.swizzle, wrap, edge_at,     \
        L.zbase, zend
    switch (L.variant) {
    case 4: jacobi3d_f64_tb<2, 16, 3><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 5: jacobi3d_f64_tb<3, 16, 3><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 6: jacobi3d_f64_tb<4, 16, 3><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 7: jacobi3d_f64_tb<2, 8, 3><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 8: jacobi3d_f64_tb<3, 8, 3><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 9: jacobi3d_f64_tb2<4, 8><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 10: jacobi3d_f64_tb2<4, 16><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 11: jacobi3d_f64_tb<1, 16, 3><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    case 12: jacobi3d_f64_tb2<1, 16><<<L.grid, L.block, 0, s>>>(LGD_TB_ARGS); break;
    default: return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "not a temporal-blocking variant");
    }
#undef LGD_TB_ARGS
    LGD_LAUNCHED("jacobi3d_f64_tb");
    return 0;
}

int launch_sweep(const double *src, double *dst, const lgdhip_grid &g, const JacobiLaunch &L,
                 hipStream_t s) {
    const int64_t sxy = g.sx * g.sy;
    const int64_t origin = lgdhip::origin_index(g);
    if (L.variant != 1) {
        switch (L.by) {
        case 1: launch_batched<1>(src, dst, g, L, s); break;
        case 2: launch_batched<2>(src, dst, g, L, s); break;
        case 8: launch_batched<8>(src, dst, g, L, s); break;
        default: launch_batched<4>(src, dst, g, L, s); break;
        }
        LGD_LAUNCHED("jacobi3d_f64_zbatch");
        return 0;
    }
    switch (L.by) {
    case 1:
        jacobi3d_f64_zmarch<1><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                         origin, L.zchunk);
        break;
    case 2:
        jacobi3d_f64_zmarch<2><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                         origin, L.zchunk);
        break;
    case 8:
        jacobi3d_f64_zmarch<8><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                         origin, L.zchunk);
        break;
    default:
        jacobi3d_f64_zmarch<4><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                         origin, L.zchunk);
        break;
    }
    LGD_LAUNCHED("jacobi3d_f64_zmarch");
    return 0;
}

} // namespace

extern "C" int lgdhip_jacobi3d_f64_plan(const lgdhip_grid *g, const lgdhip_tile *tile, int *variant, int *zchunk) {
    int rc = lgdhip::check_grid(g, 3, 8, 1, true);
    if (rc) return rc;
    LGD_REQUIRE(variant && zchunk, "NULL output pointer");
    lgdhip_tile view_tile;
    if (g->flags & LGDHIP_GRID_VIEW) {  // as lgdhip_jacobi3d_f64_run treats views
        if (tile && tb_shape(tile->variant).k > 0)
            return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "temporal blocking needs a whole grid, not a view");
        if (!tile || tile->variant == 0) {
            view_tile = tile ? *tile : lgdhip_tile{};
            view_tile.variant = kSingleSweepVariant;
            tile = &view_tile;
        }
    }
    const JacobiLaunch L = plan(*g, tile);
    *variant = L.variant;
    *zchunk = L.zchunk;
    return 0;
}

extern "C" int lgdhip_jacobi3d_f64_run(double *a, double *b, const lgdhip_grid *g, int steps,
                                       const lgdhip_tile *tile, int *result_in_b, lgdhip_stream_t stream) {
    int rc = lgdhip::check_grid(g, 3, 8, 1, true);
    if (rc) return rc;
    LGD_REQUIRE(a && b && a != b, "need two distinct device buffers");
    LGD_REQUIRE(steps >= 0, "steps must be >= 0");
    // a view's surroundings are real cells, not the edge: only the single
    // sweeps (which read neighbours from memory) apply; the default picks one
    lgdhip_tile view_tile;
    if (g->flags & LGDHIP_GRID_VIEW) {
        if (tile && tb_shape(tile->variant).k > 0)
            return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "temporal blocking needs a whole grid, not a view");
        if (!tile || tile->variant == 0) {
            view_tile = tile ? *tile : lgdhip_tile{};
            view_tile.variant = kSingleSweepVariant;
            tile = &view_tile;
        }
    }
    const JacobiLaunch L = plan(*g, tile);
    hipStream_t s = lgdhip::as_stream(stream);
    double *cur = a, *nxt = b;
    int t = 0;
    const int k = tb_shape(L.variant).k;
    if (k > 0) {
        // the blocked kernel reads memory only inside the box and its
        // one-cell halo shell (validated: halo >= 1), which holds the edge
        for (; t + k <= steps; t += k) {
            rc = launch_tb(cur, nxt, *g, L, s);
            if (rc) return rc;
            double *tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
    }
    JacobiLaunch single = L;
    if (k > 0) {
        lgdhip_tile plain = tile ? *tile : lgdhip_tile{};
        plain.variant = kSingleSweepVariant;
        plain.by = 0;
        plain.bz = 0;
        single = plan(*g, &plain);
    }
    for (; t < steps; ++t) {
        rc = launch_sweep(cur, nxt, *g, single, s);
        if (rc) return rc;
        double *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (result_in_b) *result_in_b = (cur == b) ? 1 : 0;
    return 0;
}

// Periodic axes (Torus topologies and selectively wrapped ones): the same
// K-sweep launches with wrapped tile coordinates, so a wrapped grid costs what
// a Cube grid does; leftover steps (steps mod K) are single sweeps, each after
// a periodic halo refresh.  Both buffers' halos are refreshed first, so the
// halo cells of the unwrapped axes hold `edge` whatever the caller left there.
extern "C" int lgdhip_jacobi3d_f64_run_periodic(double *a, double *b, const lgdhip_grid *g, int steps, int wrap_mask,
                                                double edge, const lgdhip_tile *tile, int *result_in_b,
                                                lgdhip_stream_t stream) {
    int rc = lgdhip::check_grid(g, 3, 8, 1, true);
    if (rc) return rc;
    LGD_REQUIRE(a && b && a != b, "need two distinct device buffers");
    LGD_REQUIRE(steps >= 0, "steps must be >= 0");
    LGD_REQUIRE(wrap_mask >= 0 && wrap_mask <= 7, "wrap_mask must be a combination of bits 1, 2, 4");
    if (g->flags & LGDHIP_GRID_VIEW)
        return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "periodic sweeps need a whole grid, not a view");
    hipStream_t s = lgdhip::as_stream(stream);
    if (steps == 0) {
        if (result_in_b) *result_in_b = 0;
        return 0;
    }
    JacobiLaunch L = plan(*g, tile);
    int k = tb_shape(L.variant).k;
    // wrapped z: a single period's correction must reach K + 3 planes past the box
    if ((wrap_mask & 4) && g->nz < 8) k = 0;
    const int64_t origin = lgdhip::origin_index(*g);
    const int64_t edge_at = !(wrap_mask & 1)   ? origin - 1
                            : !(wrap_mask & 2) ? origin - g->sx
                            : !(wrap_mask & 4) ? origin - g->sx * g->sy
                                               : origin;
    double *cur = a, *nxt = b;
    int t = 0;
    if (k > 0 && steps >= k) {
        rc = lgdhip_grid_fill_halo(a, g, &edge, wrap_mask, stream);
        if (rc) return rc;
        rc = lgdhip_grid_fill_halo(b, g, &edge, wrap_mask, stream);
        if (rc) return rc;
        for (; t + k <= steps; t += k) {
            rc = launch_tb(cur, nxt, *g, L, s, wrap_mask, edge_at);
            if (rc) return rc;
            double *tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
    }
    lgdhip_tile plain = tile ? *tile : lgdhip_tile{};
    if (k > 0 || tb_shape(L.variant).k > 0 || plain.variant == 0) {
        plain.variant = kSingleSweepVariant;
        plain.by = 0;
        plain.bz = 0;
    }
    const JacobiLaunch single = plan(*g, &plain);
    for (; t < steps; ++t) {
        rc = lgdhip_grid_fill_halo(cur, g, &edge, wrap_mask, stream);
        if (rc) return rc;
        rc = launch_sweep(cur, nxt, *g, single, s);
        if (rc) return rc;
        double *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (result_in_b) *result_in_b = (cur == b) ? 1 : 0;
    return 0;
}

// One blocked launch writing only the planes [z0, z1) of dst (all x, y),
// K sweeps of src (K of the tile's variant; the default plan's otherwise):
// the building block of a stepper that updates a slab's rim planes first,
// hands them to a copy, and then updates the inner planes.
extern "C" int lgdhip_jacobi3d_f64_run_zrange(const double *src, double *dst, const lgdhip_grid *g, int z0, int z1,
                                              const lgdhip_tile *tile, int *sweeps, lgdhip_stream_t stream) {
    int rc = lgdhip::check_grid(g, 3, 8, 1, true);
    if (rc) return rc;
    LGD_REQUIRE(src && dst && src != dst, "need two distinct device buffers");
    LGD_REQUIRE(sweeps, "NULL output pointer");
    if (g->flags & LGDHIP_GRID_VIEW)
        return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "a z range of a blocked launch needs a whole grid");
    LGD_REQUIRE(0 <= z0 && z0 < z1 && z1 <= g->nz, "need 0 <= z0 < z1 <= nz");
    JacobiLaunch L = plan(*g, tile);
    const int k = tb_shape(L.variant).k;
    if (k < 1) return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "a z range needs a temporal-blocking variant");
    if (!tile || tile->bz <= 0) L.zchunk = std::min(L.zchunk, z1 - z0);
    L.zbase = z0;
    L.zend = z1;
    L.ntiles = L.gx * L.gy * lgdhip::ceil_div(z1 - z0, L.zchunk);
    L.per = lgdhip::xcd_per(L.ntiles);
    L.grid = dim3(L.swizzle ? lgdhip::xcd_groups(L.ntiles) : L.ntiles, 1, 1);
    rc = launch_tb(src, dst, *g, L, lgdhip::as_stream(stream));
    if (rc) return rc;
    *sweeps = k;
    return 0;
}
