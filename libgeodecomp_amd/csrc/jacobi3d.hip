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

// variant ids (lgdhip_tile::variant): 0 default, 1 single-plane march,
// 2 four-plane batches, 3 eight-plane batches
constexpr int kJacobiDefaultVariant = 2;

struct JacobiLaunch {
    dim3 grid, block;
    int zchunk;
    int by;
    int variant;
};

JacobiLaunch plan(const lgdhip_grid &g, const lgdhip_tile *tile) {
    JacobiLaunch L;
    L.by = 4;
    L.zchunk = 16;
    L.variant = kJacobiDefaultVariant;
    if (tile) {
        if (tile->by == 1 || tile->by == 2 || tile->by == 4 || tile->by == 8) L.by = tile->by;
        if (tile->bz > 0) L.zchunk = tile->bz;
        if (tile->variant >= 1 && tile->variant <= 3) L.variant = tile->variant;
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
    int rc = lgdhip::check_grid(g, 3, 8, 1);
    if (rc) return rc;
    LGD_REQUIRE(a && b && a != b, "need two distinct device buffers");
    LGD_REQUIRE(steps >= 0, "steps must be >= 0");
    const JacobiLaunch L = plan(*g, tile);
    hipStream_t s = lgdhip::as_stream(stream);
    double *cur = a, *nxt = b;
    for (int t = 0; t < steps; ++t) {
        rc = launch_sweep(cur, nxt, *g, L, s);
        if (rc) return rc;
        double *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (result_in_b) *result_in_b = (cur == b) ? 1 : 0;
    return 0;
}
