// D3Q19 BGK lattice Boltzmann, lid-driven cavity, FP32 populations.
//
// Structure after the reference LBM cell (src/examples/latticeboltzmann/main.cpp:92-229):
// fused pull stream + collide, incompressible equilibrium built from the
// momentum j (not u = j/rho), weights 1/3, 1/18, 1/36, direction names and
// numbering C N E W S T B NW SW NE SE TW BW TE BE TN BN TS BS (:36-55).
// Differences, all deliberate and shared with the oracle cell:
//  * relaxation rate omega is a parameter (the reference hard-codes 1/1.7);
//  * walls are half-way bounce-back decided per link from the cell's wall
//    flags (the reference uses extra boundary cells with a State enum);
//  * the +y face may be a moving lid: a population coming back from it gets
//    6 w_i (e_i . u_lid), the moving-wall bounce-back, where the reference
//    injects a fixed momentum on two links (:206-220).
// Layout: 19 members of the padded SoA grid (one per population), in either
// member layout of the grid (whole member planes, or z slabs with the 19
// members of each z slice together: the kernels take the member stride
// `plane` and the z-slice stride `sxy` from the descriptor), and a separate
// uint8 flag grid with the same logical extents.
//
// Storage is shifted: the device holds g_i = f_i - w_i (deviation from the
// rest state).  Bounce-back and streaming are unchanged (w_i = w_opp(i)),
// rho = 1 + sum g and j = sum e g, and the equilibrium deviation is
// 3 w_i (drho/3 - j.j/2 + e.j + 1.5 (e.j)^2) with drho = sum g.  Summing
// the small deviations instead of values near 1/3 removes the systematic
// FP32 mass drift of the unshifted form (about one ulp of rho per step).
#include "lgdhip_internal.h"

#include <cstring>

#include <type_traits>

namespace {

// e_i, opposite direction, weight class (0: rest, 1: axis, 2: diagonal)
struct Dir {
    int ex, ey, ez, opp, wc;
};
// C N E W S T B NW SW NE SE TW BW TE BE TN BN TS BS
constexpr Dir kDirs[19] = {
    {0, 0, 0, 0, 0},    // C
    {0, 1, 0, 4, 1},    // N
    {1, 0, 0, 3, 1},    // E
    {-1, 0, 0, 2, 1},   // W
    {0, -1, 0, 1, 1},   // S
    {0, 0, 1, 6, 1},    // T
    {0, 0, -1, 5, 1},   // B
    {-1, 1, 0, 10, 2},  // NW
    {-1, -1, 0, 9, 2},  // SW
    {1, 1, 0, 8, 2},    // NE
    {1, -1, 0, 7, 2},   // SE
    {-1, 0, 1, 14, 2},  // TW
    {-1, 0, -1, 13, 2}, // BW
    {1, 0, 1, 12, 2},   // TE
    {1, 0, -1, 11, 2},  // BE
    {0, 1, 1, 18, 2},   // TN
    {0, 1, -1, 17, 2},  // BN
    {0, -1, 1, 16, 2},  // TS
    {0, -1, -1, 15, 2}, // BS
};

__device__ __forceinline__ constexpr float weight(int wc) {
    return wc == 0 ? 1.0f / 3.0f : (wc == 1 ? 1.0f / 18.0f : 1.0f / 36.0f);
}

// is the link feeding direction d blocked by a wall, given the cell flags?
__device__ __forceinline__ bool blocked(const Dir &d, uint32_t fl) {
    return (d.ex == 1 && (fl & LGDHIP_LBM_WALL_XLO)) || (d.ex == -1 && (fl & LGDHIP_LBM_WALL_XHI)) ||
           (d.ey == 1 && (fl & LGDHIP_LBM_WALL_YLO)) || (d.ey == -1 && (fl & LGDHIP_LBM_WALL_YHI)) ||
           (d.ez == 1 && (fl & LGDHIP_LBM_WALL_ZLO)) || (d.ez == -1 && (fl & LGDHIP_LBM_WALL_ZHI));
}

// moving lid: populations reflected from the +y face gain 6 w_i e_i.u
__device__ __forceinline__ void lid_term(float (&f)[19], uint32_t fl, float lux, float luy, float luz) {
    const bool lid = (fl & LGDHIP_LBM_WALL_YHI) && (fl & LGDHIP_LBM_LID_YHI);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const Dir d = kDirs[i];
        if (d.ey == -1) {
            const float term = 6.0f * weight(d.wc) * ((float)d.ex * lux + (float)d.ey * luy + (float)d.ez * luz);
            f[i] += lid ? term : 0.0f;
        }
    }
}

// BGK collision of the streamed (shifted) populations f into out, in the
// one order every LBM kernel here uses (bit-identical results)
__device__ __forceinline__ void collide(const float (&f)[19], float omega, float (&out)[19]) {
    float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        drho += f[i];
        if (kDirs[i].ex) jx += (float)kDirs[i].ex * f[i];
        if (kDirs[i].ey) jy += (float)kDirs[i].ey * f[i];
        if (kDirs[i].ez) jz += (float)kDirs[i].ez * f[i];
    }
    const float dterm = drho * (1.0f / 3.0f) - 0.5f * (jx * jx + jy * jy + jz * jz);
    const float keep = 1.0f - omega;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const Dir d = kDirs[i];
        const float eu = (float)d.ex * jx + (float)d.ey * jy + (float)d.ez * jz;
        const float w3 = 3.0f * weight(d.wc) * omega;
        out[i] = keep * f[i] + w3 * (dterm + eu + 1.5f * eu * eu);
    }
}

// IDX is int32_t when every element offset of the 19 planes fits 31 bits
// (then addresses are a uniform base plus a 32-bit lane offset), else int64_t.
// A work-group is a 64 x BY tile of one z plane; tiles are numbered x
// fastest, then y, then z, and work-groups map to them through the XCD
// swizzle (each XCD a contiguous range), so the cache lines that x-shifted
// pulls straddle between neighbouring tiles are fetched into one L2.
template <int BX, int BY, typename IDX, bool NTL>
__global__ __launch_bounds__(BX * BY) void lbm_d3q19_pull(const float *__restrict__ src, float *__restrict__ dst,
                                                          const uint8_t *__restrict__ flags, int nx, int ny,
                                                          int nz, IDX sx, IDX sxy, IDX plane, IDX origin,
                                                          IDX fsx, IDX fsxy, IDX forigin, float omega, float lux,
                                                          float luy, float luz, unsigned gx, unsigned gy,
                                                          unsigned ntiles, unsigned per, int swizzle) {
    // swizzle: 0 launch order, 1 one contiguous tile range per XCD, 2 rows
    // of gx tiles (one x row of the grid) per XCD, rows round-robin
    const unsigned t = swizzle == 2 ? lgdhip::xcd_group_tile(blockIdx.x, gx)
                                    : lgdhip::xcd_tile(blockIdx.x, per, swizzle == 1);
    if (t >= ntiles) return;
    const int x = (int)(t % gx) * BX + threadIdx.x;
    const int y = (int)((t / gx) % gy) * BY + threadIdx.y;
    const int z = (int)(t / (gx * gy));
    if (x >= nx || y >= ny || z >= nz) return;
    const IDX idx = origin + (IDX)z * sxy + (IDX)y * sx + x;
    const uint32_t fl = flags[forigin + (IDX)z * fsxy + (IDX)y * fsx + x];

    // all 19 source offsets first (selects, no branches), then all loads,
    // so every wave has its 19 loads in flight at once
    IDX at[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const Dir d = kDirs[i];
        const IDX pulled = (IDX)i * plane + idx - (d.ex + d.ey * sx + d.ez * sxy);
        const IDX bounced = (IDX)d.opp * plane + idx;
        at[i] = blocked(d, fl) ? bounced : pulled;
    }
    float f[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) f[i] = NTL ? __builtin_nontemporal_load(src + at[i]) : src[at[i]];
    lid_term(f, fl, lux, luy, luz);
    float out[19];
    collide(f, omega, out);
    // streamed once per sweep: keep it out of the caches' way
#pragma unroll
    for (int i = 0; i < 19; ++i) __builtin_nontemporal_store(out[i], &dst[(IDX)i * plane + idx]);
}

// Four cells per thread (x0 .. x0+3, x0 a multiple of 4): every population
// is pulled with one 16-byte load (x-shifted pulls are 4-byte aligned,
// which gfx950 serves as one dwordx4 access) and written with one 16-byte
// nontemporal store; a wave covers 256 consecutive cells of a row.  Links
// blocked by a wall are fixed up afterwards, only in lanes whose four cells
// carry wall flags.  Same arithmetic per cell, in the same order, as
// lbm_d3q19_pull (bit-identical results).
typedef float float4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float float4a __attribute__((ext_vector_type(4)));

template <int BY, typename IDX>
__global__ __launch_bounds__(64 * BY) void lbm_d3q19_pull_x4(const float *__restrict__ src, float *__restrict__ dst,
                                                             const uint8_t *__restrict__ flags, int nx, int ny,
                                                             int nz, IDX sx, IDX sxy, IDX plane, IDX origin,
                                                             IDX fsx, IDX fsxy, IDX forigin, float omega, float lux,
                                                             float luy, float luz, unsigned gx, unsigned gy,
                                                             unsigned ntiles, unsigned per, int swizzle) {
    const unsigned t = swizzle == 2 ? lgdhip::xcd_group_tile(blockIdx.x, gx)
                                    : lgdhip::xcd_tile(blockIdx.x, per, swizzle == 1);
    if (t >= ntiles) return;
    const int x0 = ((int)(t % gx) * 64 + (int)threadIdx.x) * 4;
    const int y = (int)((t / gx) % gy) * BY + threadIdx.y;
    const int z = (int)(t / (gx * gy));
    if (x0 >= nx || y >= ny || z >= nz) return;
    const IDX idx = origin + (IDX)z * sxy + (IDX)y * sx + x0;
    const uint32_t fl4 = *(const uint32_t *)(flags + forigin + (IDX)z * fsxy + (IDX)y * fsx + x0);

    float f[19][4];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const Dir d = kDirs[i];
        const float4u v = *(const float4u *)(src + ((IDX)i * plane + idx - (d.ex + d.ey * sx + d.ez * sxy)));
        f[i][0] = v.x;
        f[i][1] = v.y;
        f[i][2] = v.z;
        f[i][3] = v.w;
    }
    if (fl4) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t fl = (fl4 >> (8 * c)) & 0xffu;
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const Dir d = kDirs[i];
                if (blocked(d, fl)) f[i][c] = src[(IDX)d.opp * plane + idx + c];
            }
            const bool lid = (fl & LGDHIP_LBM_WALL_YHI) && (fl & LGDHIP_LBM_LID_YHI);
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const Dir d = kDirs[i];
                if (d.ey == -1) {
                    const float term =
                        6.0f * weight(d.wc) * ((float)d.ex * lux + (float)d.ey * luy + (float)d.ez * luz);
                    f[i][c] += lid ? term : 0.0f;
                }
            }
        }
    }
    float out[19][4];
    const float keep = 1.0f - omega;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            drho += f[i][c];
            if (kDirs[i].ex) jx += (float)kDirs[i].ex * f[i][c];
            if (kDirs[i].ey) jy += (float)kDirs[i].ey * f[i][c];
            if (kDirs[i].ez) jz += (float)kDirs[i].ez * f[i][c];
        }
        const float dterm = drho * (1.0f / 3.0f) - 0.5f * (jx * jx + jy * jy + jz * jz);
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            const Dir d = kDirs[i];
            const float eu = (float)d.ex * jx + (float)d.ey * jy + (float)d.ez * jz;
            const float w3 = 3.0f * weight(d.wc) * omega;
            out[i][c] = keep * f[i][c] + w3 * (dterm + eu + 1.5f * eu * eu);
        }
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        float4a v;
        v.x = out[i][0];
        v.y = out[i][1];
        v.z = out[i][2];
        v.w = out[i][3];
        __builtin_nontemporal_store(v, (float4a *)(dst + ((IDX)i * plane + idx)));
    }
}

// Two steps per pass (temporal blocking).  A work-group owns a tile of
// 62 x BY columns and marches a z chunk; its 64 x (BY + 2) threads cover the
// tile plus a one-cell ring.  For every plane p of the chunk (and one ring
// plane on each side) the threads compute step 1 at (x, y, p) -- pull from
// the source, bounce-back, lid, collision -- into a 3-plane LDS ring holding
// all 19 populations; then the tile's inner threads compute step 2 at
// (x, y, p - 1) by pulling from that ring (planes p - 2, p - 1, p), and
// write it.  Step-1 cells outside the box hold the halo value (the edge,
// never updated, as between single steps); cells more than one outside are
// never read.  Same per-cell arithmetic as lbm_d3q19_pull (lid_term,
// collide): two passes of this kernel equal four single steps bit for bit.
constexpr int kPull2X = 62;  // output columns per tile (64 lanes minus the ring)

template <int BY, typename IDX>
__global__ __launch_bounds__(64 * (BY + 2)) void lbm_d3q19_pull2(
    const float *__restrict__ src, float *__restrict__ dst, const uint8_t *__restrict__ flags, int nx, int ny, int nz,
    IDX sx, IDX zs, IDX ms, IDX origin, IDX fsx, IDX fzs, IDX forigin, float omega, float lux, float luy, float luz,
    unsigned gx, unsigned gy, unsigned ntiles, int zchunk) {
    constexpr int RY = BY + 2;
    __shared__ float ring[3][19][RY][64];
    const unsigned t = blockIdx.x;
    if (t >= ntiles) return;  // whole work-group, before any barrier
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = (int)(t % gx) * kPull2X - 1 + tx;
    const int y = (int)((t / gx) % gy) * BY - 1 + ty;
    const int z0 = (int)(t / (gx * gy)) * zchunk;
    const int z1 = min(z0 + zchunk, nz);
    const bool inxy = (unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny;
    // one cell outside the box in x and/or y (and not farther): halo
    const bool nearxy = x >= -1 && x <= nx && y >= -1 && y <= ny;
    const bool writer = inxy && tx >= 1 && tx <= kPull2X && ty >= 1 && ty <= BY;
    const IDX col = origin + (IDX)y * sx + x;           // (x, y, 0), valid for nearxy
    const IDX fcol = forigin + (IDX)y * fsx + x;
    // step-1 inputs of plane pp: the 19 pulled (or bounced) values of a box
    // cell, the halo values of a cell one outside, zeros farther out
    auto flag_at = [&](int pp) -> uint32_t {
        return (inxy && (unsigned)pp < (unsigned)nz) ? (uint32_t)flags[fcol + (IDX)pp * fzs] : 0u;
    };
    auto load_plane = [&](int pp, uint32_t fl, float (&v)[19]) {
        const IDX idx = col + (IDX)pp * zs;
        if (inxy && (unsigned)pp < (unsigned)nz) {
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const Dir d = kDirs[i];
                const IDX pulled = (IDX)i * ms + idx - (d.ex + d.ey * sx + d.ez * zs);
                const IDX bounced = (IDX)d.opp * ms + idx;
                v[i] = src[blocked(d, fl) ? bounced : pulled];
            }
        } else if (nearxy && pp >= -1 && pp <= nz) {
#pragma unroll
            for (int i = 0; i < 19; ++i) v[i] = src[(IDX)i * ms + idx];
        } else {
#pragma unroll
            for (int i = 0; i < 19; ++i) v[i] = 0.0f;
        }
    };
    // software pipeline: while plane p is computed, the values of plane
    // p + 1 and the flags of plane p + 2 are in flight
    uint32_t fl_cur = flag_at(z0 - 1), fl_next = flag_at(z0);
    float pv[19];
    load_plane(z0 - 1, fl_cur, pv);
    uint32_t fl_prev = 0;
    int slot = 0;                                        // ring slot of plane p
    for (int p = z0 - 1; p <= z1; ++p) {
        float f[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = pv[i];
        const uint32_t fl = fl_cur;
        if (p + 1 <= z1) load_plane(p + 1, fl_next, pv);
        const uint32_t fl_after = p + 2 <= z1 ? flag_at(p + 2) : 0u;
        // ---- step 1 at (x, y, p)
        if (inxy && (unsigned)p < (unsigned)nz) {
            lid_term(f, fl, lux, luy, luz);
            float out[19];
            collide(f, omega, out);
#pragma unroll
            for (int i = 0; i < 19; ++i) ring[slot][i][ty][tx] = out[i];
        } else {
#pragma unroll
            for (int i = 0; i < 19; ++i) ring[slot][i][ty][tx] = f[i];
        }
        __syncthreads();
        // ---- step 2 at (x, y, q = p - 1)
        const int q = p - 1;
        if (writer && q >= z0 && q < z1) {
            const int sm = slot == 0 ? 2 : slot - 1;    // plane q
            const int sl = sm == 0 ? 2 : sm - 1;        // plane q - 1
            float f[19];
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                const Dir d = kDirs[i];
                if (blocked(d, fl_prev)) {
                    f[i] = ring[sm][d.opp][ty][tx];
                } else {
                    const int s = d.ez > 0 ? sl : (d.ez < 0 ? slot : sm);
                    f[i] = ring[s][i][ty - d.ey][tx - d.ex];
                }
            }
            lid_term(f, fl_prev, lux, luy, luz);
            float out[19];
            collide(f, omega, out);
            const IDX idx = col + (IDX)q * zs;
#pragma unroll
            for (int i = 0; i < 19; ++i) __builtin_nontemporal_store(out[i], &dst[(IDX)i * ms + idx]);
        }
        fl_prev = fl;
        fl_cur = fl_next;
        fl_next = fl_after;
        __syncthreads();  // the next step 1 overwrites plane q - 1's slot
        slot = slot == 2 ? 0 : slot + 1;
    }
}

__global__ void lbm_cavity_flags_kernel(uint8_t *__restrict__ flags, int nx, int ny, int nz, int64_t sx,
                                        int64_t sxy, int64_t origin) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int z = blockIdx.z;
    if (x >= nx) return;
    uint32_t fl = 0;
    if (x == 0) fl |= LGDHIP_LBM_WALL_XLO;
    if (x == nx - 1) fl |= LGDHIP_LBM_WALL_XHI;
    if (y == 0) fl |= LGDHIP_LBM_WALL_YLO;
    if (y == ny - 1) fl |= LGDHIP_LBM_WALL_YHI | LGDHIP_LBM_LID_YHI;
    if (z == 0) fl |= LGDHIP_LBM_WALL_ZLO;
    if (z == nz - 1) fl |= LGDHIP_LBM_WALL_ZHI;
    flags[origin + (int64_t)z * sxy + (int64_t)y * sx + x] = (uint8_t)fl;
}

__global__ void lbm_moments_kernel(const float *__restrict__ f, int nx, int ny, int nz, int64_t sx, int64_t sxy,
                                   int64_t plane, int64_t origin, float *__restrict__ rho, float *__restrict__ ux,
                                   float *__restrict__ uy, float *__restrict__ uz) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int z = blockIdx.z;
    if (x >= nx) return;
    const int64_t idx = origin + (int64_t)z * sxy + (int64_t)y * sx + x;
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;  // r = drho (shifted storage)
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const float v = f[(int64_t)i * plane + idx];
        r += v;
        if (kDirs[i].ex) jx += (float)kDirs[i].ex * v;
        if (kDirs[i].ey) jy += (float)kDirs[i].ey * v;
        if (kDirs[i].ez) jz += (float)kDirs[i].ez * v;
    }
    const int64_t o = ((int64_t)z * ny + y) * nx + x;
    rho[o] = 1.0f + r;
    ux[o] = jx;
    uy[o] = jy;
    uz[o] = jz;
}

int flags_grid(const lgdhip_grid &g, lgdhip_grid *fg) {
    return lgdhip_grid_make(fg, 3, g.nx, g.ny, g.nz, g.halo, 1, 1);
}

int run_pull2(float *a, float *b, const uint8_t *flags, const lgdhip_grid &g, const lgdhip_grid &fg,
              const lgdhip_lbm_params &p, int steps, const lgdhip_tile *tile, int *result_in_b, hipStream_t s) {
    const int BY = tile->by == 2 ? 2 : 4;  // 3 x 19 x (BY + 2) x 64 floats of LDS: 58 KB or 88 KB
    const int zchunk = tile->bz > 0 ? tile->bz : 32;
    const unsigned gx = lgdhip::ceil_div(g.nx, kPull2X), gy = lgdhip::ceil_div(g.ny, BY);
    const unsigned ntiles = gx * gy * lgdhip::ceil_div(g.nz, zchunk);
    const int64_t zs = lgdhip::z_stride(g), ms = lgdhip::member_stride(g);
    const int64_t origin = lgdhip::origin_index(g), forigin = lgdhip::origin_index(fg);
    const int64_t fzs = lgdhip::z_stride(fg);
    const bool small_index =
        19 * lgdhip::plane_elems(g) < ((int64_t)1 << 31) && lgdhip::plane_elems(fg) < ((int64_t)1 << 31);
    float *cur = a, *nxt = b;
    int t = 0;
    for (; t + 2 <= steps; t += 2) {
        // two steps: cur -> nxt (the intermediate step lives in LDS only)
#define LGD_PULL2(B, I)                                                                                          \
    lbm_d3q19_pull2<B, I><<<ntiles, dim3(64, B + 2), 0, s>>>(                                                    \
        cur, nxt, flags, (int)g.nx, (int)g.ny, (int)g.nz, (I)g.sx, (I)zs, (I)ms, (I)origin, (I)fg.sx, (I)fzs,    \
        (I)forigin, p.omega, p.lid_ux, p.lid_uy, p.lid_uz, gx, gy, ntiles, zchunk)
        if (small_index) {
            if (BY == 2) LGD_PULL2(2, int32_t); else LGD_PULL2(4, int32_t);
        } else {
            if (BY == 2) LGD_PULL2(2, int64_t); else LGD_PULL2(4, int64_t);
        }
#undef LGD_PULL2
        LGD_LAUNCHED("lbm_d3q19_pull2");
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (t < steps) {
        int in_b = 0;
        lgdhip_tile single;
        std::memset(&single, 0, sizeof(single));
        const int rc = lgdhip_lbm_d3q19_f32_run(cur, nxt, flags, &g, &p, 1, &single, &in_b, s);
        if (rc) return rc;
        if (in_b) {
            float *tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
    }
    if (result_in_b) *result_in_b = (cur == b) ? 1 : 0;
    return 0;
}

int run_x4(float *a, float *b, const uint8_t *flags, const lgdhip_grid &g, const lgdhip_grid &fg,
           const lgdhip_lbm_params &p, int steps, const lgdhip_tile *tile, int swizzle, int *result_in_b,
           hipStream_t s) {
    const int BY = (tile->by == 2 || tile->by == 4) ? tile->by : 1;
    const unsigned gx = lgdhip::ceil_div(g.nx, 256), gy = lgdhip::ceil_div(g.ny, BY);
    const unsigned ntiles = gx * gy * (unsigned)g.nz;
    const unsigned per = lgdhip::xcd_per(ntiles);
    const dim3 grid(swizzle == 2 ? lgdhip::xcd_group_groups(ntiles, gx)
                                 : (swizzle ? lgdhip::xcd_groups(ntiles) : ntiles));
    const dim3 block(64, BY);
    // z-slice and member strides of the grid's layout
    const int64_t sxy = lgdhip::z_stride(g);
    const int64_t plane = lgdhip::member_stride(g);
    const int64_t origin = lgdhip::origin_index(g);
    const int64_t forigin = lgdhip::origin_index(fg);
    const bool small_index =
        19 * lgdhip::plane_elems(g) < ((int64_t)1 << 31) && lgdhip::plane_elems(fg) < ((int64_t)1 << 31);
    float *cur = a, *nxt = b;
    auto launch = [&](auto by_tag) {
        constexpr int B = decltype(by_tag)::value;
        if (small_index) {
            lbm_d3q19_pull_x4<B, int32_t><<<grid, block, 0, s>>>(
                cur, nxt, flags, (int)g.nx, (int)g.ny, (int)g.nz, (int32_t)g.sx, (int32_t)sxy, (int32_t)plane,
                (int32_t)origin, (int32_t)fg.sx, (int32_t)(fg.sx * fg.sy), (int32_t)forigin, p.omega, p.lid_ux,
                p.lid_uy, p.lid_uz, gx, gy, ntiles, per, swizzle);
        } else {
            lbm_d3q19_pull_x4<B, int64_t><<<grid, block, 0, s>>>(
                cur, nxt, flags, (int)g.nx, (int)g.ny, (int)g.nz, g.sx, sxy, plane, origin, fg.sx, fg.sx * fg.sy,
                forigin, p.omega, p.lid_ux, p.lid_uy, p.lid_uz, gx, gy, ntiles, per, swizzle);
        }
    };
    for (int t = 0; t < steps; ++t) {
        switch (BY) {
        case 2: launch(std::integral_constant<int, 2>{}); break;
        case 4: launch(std::integral_constant<int, 4>{}); break;
        default: launch(std::integral_constant<int, 1>{}); break;
        }
        LGD_LAUNCHED("lbm_d3q19_pull_x4");
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (result_in_b) *result_in_b = (cur == b) ? 1 : 0;
    return 0;
}

} // namespace

extern "C" {

int lgdhip_lbm_d3q19_f32_run(float *a, float *b, const uint8_t *flags, const lgdhip_grid *g,
                             const lgdhip_lbm_params *p, int steps, const lgdhip_tile *tile, int *result_in_b,
                             lgdhip_stream_t stream) {
    int rc = lgdhip::check_grid(g, 3, 4, 19);
    if (rc) return rc;
    LGD_REQUIRE(a && b && a != b && flags && p, "NULL or aliased argument");
    LGD_REQUIRE(steps >= 0, "steps must be >= 0");
    LGD_REQUIRE(g->nz <= 65535, "nz must be <= 65535");
    lgdhip_grid fg;
    rc = flags_grid(*g, &fg);
    if (rc) return rc;
    // work-group: 64 x BY cells, or (bx = 256) 256 x 1: a whole 256-cell row
    // segment in one work-group, so lines straddled by the x-shifted pulls
    // of its four waves are fetched once into that CU's caches
    const int BX = (tile && tile->bx == 256) ? 256 : 64;
    const int BY = BX == 256 ? 1 : ((tile && (tile->by == 1 || tile->by == 2 || tile->by == 8)) ? tile->by : 4);
    // tile order: LGDHIP_TILE_XCD_SWIZZLE one range per XCD,
    // LGDHIP_TILE_XCD_ROWS x rows of tiles round-robin over the XCDs,
    // default (or LGDHIP_TILE_NO_XCD_SWIZZLE) launch order
    int swizzle = lgdhip::xcd_swizzle(tile, false);
    if (tile && (tile->flags & LGDHIP_TILE_XCD_ROWS) && !(tile->flags & LGDHIP_TILE_NO_XCD_SWIZZLE)) swizzle = 2;
    const bool ntl = tile && tile->variant == 1;
    // variant 3: two steps per pass (pairs), an odd remainder single-stepped
    if (tile && tile->variant == 3) {
        return run_pull2(a, b, flags, *g, fg, *p, steps, tile, result_in_b, lgdhip::as_stream(stream));
    }
    // variant 2: four cells per thread, 64 x BY threads = 256 x BY cells
    const bool x4 = tile && tile->variant == 2;
    if (x4) {
        if (g->nx % 4 != 0) return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "variant 2 needs nx % 4 == 0");
        return run_x4(a, b, flags, *g, fg, *p, steps, tile, swizzle, result_in_b, lgdhip::as_stream(stream));
    }
    const unsigned gx = lgdhip::ceil_div(g->nx, BX), gy = lgdhip::ceil_div(g->ny, BY);
    const unsigned ntiles = gx * gy * (unsigned)g->nz;
    const unsigned per = lgdhip::xcd_per(ntiles);
    const dim3 grid(swizzle == 2 ? lgdhip::xcd_group_groups(ntiles, gx)
                                 : (swizzle ? lgdhip::xcd_groups(ntiles) : ntiles));
    const dim3 block(BX, BY);
    // z-slice and member strides of the grid's layout
    const int64_t sxy = lgdhip::z_stride(*g);
    const int64_t plane = lgdhip::member_stride(*g);
    const int64_t origin = lgdhip::origin_index(*g);
    const int64_t forigin = lgdhip::origin_index(fg);
    // 32-bit element offsets when all 19 members (and the flag grid) fit 31 bits
    const bool small_index =
        19 * lgdhip::plane_elems(*g) < ((int64_t)1 << 31) && lgdhip::plane_elems(fg) < ((int64_t)1 << 31);
    hipStream_t s = lgdhip::as_stream(stream);
    float *cur = a, *nxt = b;
    auto launch_nt = [&](auto bx_tag, auto by_tag, auto nt_tag, const float *src, float *dst) {
        constexpr int X = decltype(bx_tag)::value;
        constexpr int B = decltype(by_tag)::value;
        constexpr bool NT = decltype(nt_tag)::value;
        if (small_index) {
            lbm_d3q19_pull<X, B, int32_t, NT><<<grid, block, 0, s>>>(
                src, dst, flags, (int)g->nx, (int)g->ny, (int)g->nz, (int32_t)g->sx, (int32_t)sxy, (int32_t)plane,
                (int32_t)origin, (int32_t)fg.sx, (int32_t)(fg.sx * fg.sy), (int32_t)forigin, p->omega, p->lid_ux,
                p->lid_uy, p->lid_uz, gx, gy, ntiles, per, swizzle);
        } else {
            lbm_d3q19_pull<X, B, int64_t, NT><<<grid, block, 0, s>>>(src, dst, flags, (int)g->nx, (int)g->ny, (int)g->nz,
                                                              g->sx, sxy, plane, origin, fg.sx, fg.sx * fg.sy,
                                                              forigin, p->omega, p->lid_ux, p->lid_uy, p->lid_uz, gx,
                                                              gy, ntiles, per, swizzle);
        }
    };
    using I64 = std::integral_constant<int, 64>;
    auto launch = [&](auto bx_tag, auto by_tag, const float *src, float *dst) {
        if (ntl)
            launch_nt(bx_tag, by_tag, std::true_type{}, src, dst);
        else
            launch_nt(bx_tag, by_tag, std::false_type{}, src, dst);
    };
    for (int t = 0; t < steps; ++t) {
        if (BX == 256) {
            launch(std::integral_constant<int, 256>{}, std::integral_constant<int, 1>{}, cur, nxt);
        } else {
            switch (BY) {
            case 1: launch(I64{}, std::integral_constant<int, 1>{}, cur, nxt); break;
            case 2: launch(I64{}, std::integral_constant<int, 2>{}, cur, nxt); break;
            case 8: launch(I64{}, std::integral_constant<int, 8>{}, cur, nxt); break;
            default: launch(I64{}, std::integral_constant<int, 4>{}, cur, nxt); break;
            }
        }
        LGD_LAUNCHED("lbm_d3q19_pull");
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (result_in_b) *result_in_b = (cur == b) ? 1 : 0;
    return 0;
}

int lgdhip_lbm_cavity_flags(uint8_t *flags, const lgdhip_grid *fg, lgdhip_stream_t stream) {
    int rc = lgdhip::check_grid(fg, 3, 1, 1);
    if (rc) return rc;
    LGD_REQUIRE(flags, "NULL pointer");
    LGD_REQUIRE(fg->ny <= 65535 && fg->nz <= 65535, "ny, nz must be <= 65535");
    dim3 grid(lgdhip::ceil_div(fg->nx, 256), (unsigned)fg->ny, (unsigned)fg->nz);
    lbm_cavity_flags_kernel<<<grid, 256, 0, lgdhip::as_stream(stream)>>>(
        flags, (int)fg->nx, (int)fg->ny, (int)fg->nz, fg->sx, fg->sx * fg->sy, lgdhip::origin_index(*fg));
    LGD_LAUNCHED("lbm_cavity_flags");
    return 0;
}

int lgdhip_lbm_d3q19_f32_moments(const float *f, const lgdhip_grid *g, float *rho, float *ux, float *uy, float *uz,
                                 lgdhip_stream_t stream) {
    int rc = lgdhip::check_grid(g, 3, 4, 19);
    if (rc) return rc;
    LGD_REQUIRE(f && rho && ux && uy && uz, "NULL pointer");
    LGD_REQUIRE(g->ny <= 65535 && g->nz <= 65535, "ny, nz must be <= 65535");
    dim3 grid(lgdhip::ceil_div(g->nx, 256), (unsigned)g->ny, (unsigned)g->nz);
    lbm_moments_kernel<<<grid, 256, 0, lgdhip::as_stream(stream)>>>(
        f, (int)g->nx, (int)g->ny, (int)g->nz, g->sx, lgdhip::z_stride(*g), lgdhip::member_stride(*g),
        lgdhip::origin_index(*g), rho, ux, uy, uz);
    LGD_LAUNCHED("lbm_moments");
    return 0;
}

} // extern "C"
