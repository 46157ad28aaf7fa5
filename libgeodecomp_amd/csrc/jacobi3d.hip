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
            dst[i + (int64_t)k * sxy] = s * (1.0 / 7.0);
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
                                                           unsigned ntiles, unsigned per, int swizzle) {
    constexpr int OX = 64 - 2 * K, OY = TY - 2 * K;
    static_assert(PF == 3, "level-0 ring of 6: a window of 3 planes plus 3 in flight");
    __shared__ double lds[2][K][TY][64];
    const unsigned t = lgdhip::xcd_tile(blockIdx.x, per, swizzle != 0);
    if (t >= ntiles) return;  // whole work-group, before any barrier
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = (int)(t % gx) * OX - K + tx;
    const int y = (int)((t / gx) % gy) * OY - K + ty;
    const int z0 = (int)(t / (gx * gy)) * zchunk;
    const int z1 = min(z0 + zchunk, nz);
    const bool inxy = (unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny;
    const bool writer = inxy && tx >= K && tx < 64 - K && ty >= K && ty < TY - K;
    const double edge = src[origin - 1];  // the Cube edge value held by every halo cell
    const int64_t col = origin + (int64_t)y * sx + x;
    const int yd = ty > 0 ? ty - 1 : 0;
    const int yu = ty < TY - 1 ? ty + 1 : TY - 1;

    // level-0 loads are unconditional: coordinates outside the box are
    // clamped into the one-cell halo shell, which holds the edge value, so
    // no select or branch guards them and the compiler's wait counts let
    // PF planes stay in flight
    const int64_t colh = origin + (int64_t)min(max(y, -1), ny) * sx + min(max(x, -1), nx);
    auto load0 = [&](int p) -> double { return src[colh + (int64_t)min(max(p, -1), nz) * sxy]; };

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
            const double v = (inxy && (unsigned)q < (unsigned)nz) ? s * (1.0 / 7.0) : edge;
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

// variant ids (lgdhip_tile::variant): 0 default, 1 single-plane march,
// 2 four-plane batches, 3 eight-plane batches; temporal blocking (K sweeps
// per launch, tile 64 x TY): 4 K=2 TY=16, 5 K=3 TY=16, 6 K=4 TY=16,
// 7 K=2 TY=8, 8 K=3 TY=8.  The default blocks 4 sweeps per launch (fastest
// on 128^3 and 512^3 on MI355X); step counts not divisible by K finish with
// single sweeps of kSingleSweepVariant.
constexpr int kJacobiDefaultVariant = 6;
constexpr int kSingleSweepVariant = 2;

struct TbShape {
    int k, ty;
};
TbShape tb_shape(int variant) {
    switch (variant) {
    case 4: return {2, 16};
    case 5: return {3, 16};
    case 6: return {4, 16};
    case 7: return {2, 8};
    case 8: return {3, 8};
    default: return {0, 0};
    }
}

struct JacobiLaunch {
    dim3 grid, block;
    int zchunk;
    int by;
    int variant;
    unsigned gx, gy, ntiles, per;  // temporal blocking: tile grid and XCD swizzle
    int swizzle;
};

JacobiLaunch plan(const lgdhip_grid &g, const lgdhip_tile *tile) {
    JacobiLaunch L;
    L.by = 4;
    L.zchunk = 16;
    L.variant = kJacobiDefaultVariant;
    if (tile) {
        if (tile->by == 1 || tile->by == 2 || tile->by == 4 || tile->by == 8) L.by = tile->by;
        if (tile->bz > 0) L.zchunk = tile->bz;
        if (tile->variant >= 1 && tile->variant <= 8) L.variant = tile->variant;
    }
    const TbShape tb = tb_shape(L.variant);
    if (tb.k > 0) {
        // temporal blocking: tiles of (64-2K) x (TY-2K) output columns; a z
        // chunk of C planes costs its work-group C full pipeline steps plus
        // a 2K-step ramp of about K - 1 steps' work.  Count the launch as
        // ceil(work-groups / CUs) rounds of C + K planes and take the chunk
        // (any C up to 64, the measured best at 512^3) minimising it: at
        // 128^3 that is 5 chunks of 26 planes, 240 work-groups.
        const unsigned gx = lgdhip::ceil_div(g.nx, 64 - 2 * tb.k);
        const unsigned gy = lgdhip::ceil_div(g.ny, tb.ty - 2 * tb.k);
        if (!tile || tile->bz <= 0) {
            uint64_t best = ~(uint64_t)0;
            for (int64_t n = lgdhip::ceil_div(g.nz, 64); n <= g.nz; ++n) {
                const int c = (int)lgdhip::ceil_div(g.nz, n);
                if (lgdhip::ceil_div(g.nz, c) != n) continue;  // same chunk as a smaller n
                const uint64_t groups = (uint64_t)gx * gy * n;
                const uint64_t cost = (groups + lgdhip::kCUs - 1) / lgdhip::kCUs * (uint64_t)(c + tb.k);
                if (cost < best) {
                    best = cost;
                    L.zchunk = c;
                }
            }
        }
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
    // aim for several thousand work-groups so all 256 CUs stay busy even on
    // small cubes: shrink the z chunk until the grid is large enough
    const unsigned gx = lgdhip::ceil_div(g.nx, kWaveX);
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
    if (L.variant == 3)
        jacobi3d_f64_zbatch<BY, 8><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                             origin, L.zchunk);
    else
        jacobi3d_f64_zbatch<BY, 4><<<L.grid, L.block, 0, s>>>(src, dst, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy,
                                                             origin, L.zchunk);
}

// one launch of the temporal-blocked kernel: K sweeps src -> dst
int launch_tb(const double *src, double *dst, const lgdhip_grid &g, const JacobiLaunch &L, hipStream_t s) {
    const int64_t sxy = g.sx * g.sy;
    const int64_t origin = lgdhip::origin_index(g);
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    switch (L.variant) {
    case 4:
        jacobi3d_f64_tb<2, 16, 3><<<L.grid, L.block, 0, s>>>(src, dst, nx, ny, nz, g.sx, sxy, origin, L.zchunk, L.gx,
                                                            L.gy, L.ntiles, L.per, L.swizzle);
        break;
    case 5:
        jacobi3d_f64_tb<3, 16, 3><<<L.grid, L.block, 0, s>>>(src, dst, nx, ny, nz, g.sx, sxy, origin, L.zchunk, L.gx,
                                                            L.gy, L.ntiles, L.per, L.swizzle);
        break;
    case 6:
        jacobi3d_f64_tb<4, 16, 3><<<L.grid, L.block, 0, s>>>(src, dst, nx, ny, nz, g.sx, sxy, origin, L.zchunk, L.gx,
                                                            L.gy, L.ntiles, L.per, L.swizzle);
        break;
    case 7:
        jacobi3d_f64_tb<2, 8, 3><<<L.grid, L.block, 0, s>>>(src, dst, nx, ny, nz, g.sx, sxy, origin, L.zchunk, L.gx,
                                                            L.gy, L.ntiles, L.per, L.swizzle);
        break;
    case 8:
        jacobi3d_f64_tb<3, 8, 3><<<L.grid, L.block, 0, s>>>(src, dst, nx, ny, nz, g.sx, sxy, origin, L.zchunk, L.gx,
                                                            L.gy, L.ntiles, L.per, L.swizzle);
        break;
    default: return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "not a temporal-blocking variant");
    }
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
