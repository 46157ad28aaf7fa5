// Batched ensemble of independent 3D Poisson problems solved by Jacobi.
//
// Problem p: n^3 interior unknowns, discrete Poisson equation
//     6 u_c - sum_{6 neighbours} u_nb = f_c
// (h = 1), constant Dirichlet value per face outside the cube.  Jacobi:
//     u_c <- (sum_nb u_nb + f_c) / 6.
// The reference has no counterpart (no source term, residual or
// batching); the sweep is the Jacobi cell of
// src/testbed/performancetests/main.cpp:1194-1205 with a source term.
//
// Precision scheme: FP32 storage of u cannot reach a 1e-6 relative
// residual (the rounding floor of u is ~1e-4 of the initial residual for
// these sizes), so the solver is a mixed-precision refinement:
//   u (FP64) -> r = f + sum_nb u - 6u in FP64, stored FP32, with the FP64
//   residual norm;  e (FP32) <- m Jacobi sweeps on A e = r from e = 0
//   (12 bytes per cell and sweep: read e, read r, write e);  u += e.
// Jacobi on the correction is the same linear iteration as Jacobi on u, so
// the sweep count to a tolerance matches plain FP64 Jacobi; only the
// arithmetic of the hot loop is FP32.
//
// Layout: each problem occupies a padded (n+2)^3 block, x fastest, at its
// element offset in the pools (u: double, f/e/r: float).  u's halo holds the
// face values; e's halo stays zero.  All kernels walk one list of tiles
// (problem, x0, y0, z0) covering every active problem, so one launch sweeps
// the whole ensemble; norms are reduced deterministically (per-tile FP64
// partials, then per-problem sums in tile order).
#include "lgdhip_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

constexpr int kTX = 32;  // tile width (x), one 128-byte FP32 row segment
constexpr int kTY = 8;   // tile height (y)
constexpr int kTZ = 16;  // z planes marched per tile
constexpr int kThreads = kTX * kTY;

struct ProblemDev {
    int32_t n;
    int32_t pad;
    int64_t base;  // element offset of the padded block
};

struct Tile {
    int32_t p, x0, y0, z0;
};

__device__ __forceinline__ double block_sum(double v, double *lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kThreads / 64; ++w) s += lds[w];
    }
    return s;
}

// r = f + sum_nb u - 6 u (FP64), r32 = (float) r, per-tile sum of r^2
__global__ __launch_bounds__(kThreads) void poisson_residual_f64(const double *__restrict__ u,
                                                                 const float *__restrict__ f,
                                                                 float *__restrict__ r32, const Tile *__restrict__ tiles,
                                                                 const ProblemDev *__restrict__ probs,
                                                                 double *__restrict__ partial) {
    __shared__ double lds[kThreads / 64];
    const Tile t = tiles[blockIdx.x];
    const ProblemDev P = probs[t.p];
    const int n = P.n;
    const int x = t.x0 + (int)(threadIdx.x % kTX);
    const int y = t.y0 + (int)(threadIdx.x / kTX);
    const int64_t sx = n + 2, sxy = sx * sx;
    double acc = 0.0;
    if (x < n && y < n) {
        const int z1 = min(t.z0 + kTZ, n);
        int64_t i = P.base + ((int64_t)(t.z0 + 1) * sx + (y + 1)) * sx + (x + 1);
        for (int z = t.z0; z < z1; ++z, i += sxy) {
            double s = u[i - sxy];
            s += u[i - sx];
            s += u[i - 1];
            s += u[i + 1];
            s += u[i + sx];
            s += u[i + sxy];
            const double r = (double)f[i] + s - 6.0 * u[i];
            r32[i] = (float)r;
            acc += r * r;
        }
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// one FP32 Jacobi sweep on A e = r:  e' = (e[z-1] + e[y-1] + e[x-1] + e[x+1]
// + e[y+1] + e[z+1] + r) * (1/6), summed in that order; when `partial` is
// given, also the per-tile sum of the FP64 square of the residual of the
// input iterate, r + sum_nb e - 6 e (= 6 (e' - e) up to rounding).
// DAMPED: weighted Jacobi e' = e + omega (J(e) - e), which damps the
// highest-frequency error modes by |1 - 2 omega| per sweep instead of
// plain Jacobi's cos(pi/(n+1)) (used to smooth a predicted initial guess).
template <bool DAMPED>
__global__ __launch_bounds__(kThreads) void poisson_sweep_f32(const float *__restrict__ e, float *__restrict__ e2,
                                                              const float *__restrict__ r,
                                                              const Tile *__restrict__ tiles,
                                                              const ProblemDev *__restrict__ probs,
                                                              double *__restrict__ partial, float omega) {
    __shared__ double lds[kThreads / 64];
    const Tile t = tiles[blockIdx.x];
    const ProblemDev P = probs[t.p];
    const int n = P.n;
    const int x = t.x0 + (int)(threadIdx.x % kTX);
    const int y = t.y0 + (int)(threadIdx.x / kTX);
    const int64_t sx = n + 2, sxy = sx * sx;
    double acc = 0.0;
    if (x < n && y < n) {
        const int z1 = min(t.z0 + kTZ, n);
        int64_t i = P.base + ((int64_t)(t.z0 + 1) * sx + (y + 1)) * sx + (x + 1);
        float below = e[i - sxy];
        float center = e[i];
        for (int z = t.z0; z < z1; ++z, i += sxy) {
            const float above = e[i + sxy];
            float s = below;
            s += e[i - sx];
            s += e[i - 1];
            s += e[i + 1];
            s += e[i + sx];
            s += above;
            s += r[i];
            if (DAMPED)  // single roundings, the same in every sweep kernel
                e2[i] = __builtin_fmaf(omega, __builtin_fmaf(s, 1.0f / 6.0f, -center), center);
            else
                e2[i] = s * (1.0f / 6.0f);
            if (partial) {
                const double rr = (double)s - 6.0 * (double)center;
                acc += rr * rr;
            }
            below = center;
            center = above;
        }
    }
    if (partial) {
        const double s = block_sum(acc, lds);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// value of lane-1 / lane+1 (DPP wave_shr:1 / wave_shl:1; the wave's end
// lanes read 0 and only feed overlap cells)
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}

// K FP32 Jacobi sweeps on A e = r per launch (temporal blocking).  A
// work-group is a 64 x TY tile of (x, y) columns of one problem with a K-cell
// overlap ring (it writes (64-2K) x (TY-2K) columns; tile.x0/y0 is the
// origin of that written region) and marches a z chunk as a pipeline: when
// plane p of the input arrives, level l is computed at plane p-l from level
// l-1 (3-plane register window per level, x neighbours by DPP, y neighbours
// through a double-buffered LDS plane per level, r from a register ring;
// the march is unrolled by 6 so every ring rotates by compile-time slot roles).
// Cells outside the problem hold 0 at every level (the correction's zero
// Dirichlet halo) and every cell sums in poisson_sweep_f32's order, so one
// launch equals K launches of that kernel bit for bit.  With `partial`, the
// last level adds the FP64 square of the residual of its input iterate over
// the written cells into one partial per tile.
template <int K, int TY, bool DAMPED>
__global__ __launch_bounds__(64 * TY) void poisson_sweep_tb(const float *__restrict__ e, float *__restrict__ e2,
                                                            const float *__restrict__ r,
                                                            const Tile *__restrict__ tiles,
                                                            const ProblemDev *__restrict__ probs,
                                                            double *__restrict__ partial, int zchunk,
                                                            unsigned ntiles, unsigned per, float omega) {
    constexpr int PF = 3;  // e planes in flight
    static_assert(K + 2 <= 6, "r ring of 6 slots holds planes p-K .. p+1");
    __shared__ float lds[2][K][TY][64];
    __shared__ double red[TY];
    // XCD c runs the contiguous tile range [c*per, (c+1)*per), so tiles whose
    // overlap rings share lines run on one L2 at about the same time
    const unsigned tid = lgdhip::xcd_tile(blockIdx.x, per, true);
    if (tid >= ntiles) return;  // whole work-group, before any barrier
    const Tile t = tiles[tid];
    const ProblemDev P = probs[t.p];
    const int n = P.n;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = t.x0 - K + tx, y = t.y0 - K + ty;
    const int z0 = t.z0, z1 = min(t.z0 + zchunk, n);
    const bool inxy = (unsigned)x < (unsigned)n && (unsigned)y < (unsigned)n;
    const bool writer = inxy && tx >= K && tx < 64 - K && ty >= K && ty < TY - K;
    const int64_t sx = n + 2, sxy = sx * sx;
    const int64_t col = P.base + (int64_t)(y + 1) * sx + (x + 1);  // padded column, plane index = z + 1
    const int yd = ty > 0 ? ty - 1 : 0;
    const int yu = ty < TY - 1 ? ty + 1 : TY - 1;
    // loads are unconditional: coordinates outside the problem are clamped
    // into its one-cell halo shell, where e is zero; r there only feeds cells
    // that are masked to zero.  No guard, so PF planes stay in flight.
    const int64_t colh = P.base + (int64_t)(min(max(y, -1), n) + 1) * sx + (min(max(x, -1), n) + 1);
    auto at = [&](int p) { return colh + (int64_t)(min(max(p, -1), n) + 1) * sxy; };

    // ev: e planes, a ring of 6 (plane p in slot (p - pbeg) mod 6: the window
    // p-2 .. p plus p+1 .. p+3 in flight); rv: r planes p-K .. p+1, a ring
    // of 6; w[l] (l >= 1): level l's last three planes.  Slot roles follow
    // the step index mod 6, a compile-time constant once the march is
    // unrolled by 6, so no value moves between registers.
    float ev[6], rv[6];
    float w[K][3];
#pragma unroll
    for (int l = 0; l < K; ++l) w[l][0] = w[l][1] = w[l][2] = 0.0f;
#pragma unroll
    for (int l = 0; l < K; ++l) lds[0][l][ty][tx] = 0.0f;
    const int pbeg = z0 - K;  // input planes pbeg .. z1 + K - 1: 2K ramp steps, then z1 - z0 steps
    ev[4] = ev[5] = 0.0f;
#pragma unroll
    for (int d = 0; d < PF; ++d) ev[d] = e[at(pbeg + d)];
    rv[0] = r[at(pbeg)];
    __syncthreads();

    double acc = 0.0;
    // step j at input plane p = pbeg + j; JM = j mod 6; JR = j during the
    // ramp, 2K after.  Level l's plane p - l is read by a written plane only
    // when j >= 2l, and level l's LDS copy only when j >= 2l + 1: the ramp
    // skips the rest (no later step reads those slots).
    auto step = [&](auto jm_c, auto jr_c, int p) {
        constexpr int JM = decltype(jm_c)::value;
        constexpr int JR = decltype(jr_c)::value;
        constexpr int NEW = JM % 3, OLD = (JM + 1) % 3, MID = (JM + 2) % 3;  // roles after this step's insert
        constexpr int B = JM & 1;                                             // LDS buffer holding step j-1's planes
        ev[(JM + PF) % 6] = e[at(p + PF)];
        rv[(JM + 1) % 6] = r[at(p + 1)];
#pragma unroll
        for (int l = 1; l <= K; ++l) {
            if (2 * l > JR) continue;
            const int q = p - l;
            const float a = l == 1 ? ev[(JM + 4) % 6] : w[l - 1][OLD];
            const float c = l == 1 ? ev[(JM + 5) % 6] : w[l - 1][MID];
            const float b = l == 1 ? ev[JM] : w[l - 1][NEW];
            float sum = a;                    // ( 0,  0, -1)
            sum += lds[B][l - 1][yd][tx];     // ( 0, -1,  0)
            sum += lane_prev(c);              // (-1,  0,  0)
            sum += lane_next(c);              // ( 1,  0,  0)
            sum += lds[B][l - 1][yu][tx];     // ( 0,  1,  0)
            sum += b;                         // ( 0,  0,  1)
            sum += rv[(JM + 6 - l) % 6];      // r at plane q
            const bool live = inxy && (unsigned)q < (unsigned)n;
            const float jv = DAMPED ? __builtin_fmaf(omega, __builtin_fmaf(sum, 1.0f / 6.0f, -c), c)
                                    : sum * (1.0f / 6.0f);
            const float v = live ? jv : 0.0f;
            if (l < K) {
                w[l][NEW] = v;
            } else if (writer) {  // after the ramp q = p - K runs over z0 .. z1 - 1
                e2[col + (int64_t)(q + 1) * sxy] = v;
                if (partial) {
                    const double rr = (double)sum - 6.0 * (double)c;
                    acc += rr * rr;
                }
            }
        }
        // publish level l at plane p - l for the next step's y neighbours
        if (JR >= 1) lds[B ^ 1][0][ty][tx] = ev[JM];
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
    if (partial) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (tx == 0) red[ty] = acc;
        __syncthreads();
        if (tx == 0 && ty == 0) {
            double sum = 0.0;
            for (int i = 0; i < TY; ++i) sum += red[i];
            partial[tid] = sum;
        }
    }
}

// The same K-sweep pipeline with R adjacent rows per thread (R = 2, 3): a
// 64 x TY work-group holds a 64 x R*TY tile of one problem (written region
// (64-2K) x (R*TY-2K); tile.x0/y0 its origin).  A thread's rows are each
// other's y neighbours in registers, LDS carries only each thread's first
// and last row for the neighbouring threads, and the taller tile recomputes
// a smaller overlap ring.  Cells, sum order, rings, ramp and residual
// partials as in poisson_sweep_tb: bit-exact with K launches of
// poisson_sweep_f32 (DAMPED: of its weighted form).
template <int K, int TY, int R, bool DAMPED>
__global__ __launch_bounds__(64 * TY) void poisson_sweep_tb2(const float *__restrict__ e, float *__restrict__ e2,
                                                             const float *__restrict__ r,
                                                             const Tile *__restrict__ tiles,
                                                             const ProblemDev *__restrict__ probs,
                                                             double *__restrict__ partial, int zchunk,
                                                             unsigned ntiles, unsigned per, float omega) {
    constexpr int PF = 3;
    constexpr int CY = R * TY;
    static_assert(K + 2 <= 6, "r ring of 6 slots holds planes p-K .. p+1");
    static_assert(R >= 2 && R <= 4, "rows per thread");
    __shared__ float lds[2][K][TY][2][64];  // [buffer][level][thread row][first, last row][lane]
    __shared__ double red[TY];
    const unsigned tid = lgdhip::xcd_tile(blockIdx.x, per, true);
    if (tid >= ntiles) return;  // whole work-group, before any barrier
    const Tile t = tiles[tid];
    const ProblemDev P = probs[t.p];
    const int n = P.n;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = t.x0 - K + tx, y0 = t.y0 - K + R * ty;  // rows y0 .. y0 + R - 1
    const int z0 = t.z0, z1 = min(t.z0 + zchunk, n);
    const bool xin = (unsigned)x < (unsigned)n;
    const bool xw = tx >= K && tx < 64 - K;
    const int64_t sx = n + 2, sxy = sx * sx;
    // unconditional loads clamped into the problem's halo shell (e is zero
    // there; r there only feeds cells masked to zero)
    const int xc = min(max(x, -1), n) + 1;
    bool inxy[R], writer[R];
    int64_t col[R], colh[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int yq = y0 + q, tr = R * ty + q;
        inxy[q] = xin && (unsigned)yq < (unsigned)n;
        writer[q] = inxy[q] && xw && tr >= K && tr < CY - K;
        col[q] = P.base + (int64_t)(yq + 1) * sx + (x + 1);
        colh[q] = P.base + (int64_t)(min(max(yq, -1), n) + 1) * sx + xc;
    }
    auto at = [&](int q, int p) { return colh[q] + (int64_t)(min(max(p, -1), n) + 1) * sxy; };
    const int tyd = ty > 0 ? ty - 1 : 0;
    const int tyu = ty < TY - 1 ? ty + 1 : TY - 1;

    float ev[R][6], rv[R][6];
    float w[K][R][3];
#pragma unroll
    for (int l = 0; l < K; ++l)
#pragma unroll
        for (int q = 0; q < R; ++q) w[l][q][0] = w[l][q][1] = w[l][q][2] = 0.0f;
#pragma unroll
    for (int l = 0; l < K; ++l) lds[0][l][ty][0][tx] = lds[0][l][ty][1][tx] = 0.0f;
    const int pbeg = z0 - K;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        ev[q][4] = ev[q][5] = 0.0f;
#pragma unroll
        for (int d = 0; d < PF; ++d) ev[q][d] = e[at(q, pbeg + d)];
        rv[q][0] = r[at(q, pbeg)];
    }
    __syncthreads();

    double acc = 0.0;
    auto step = [&](auto jm_c, auto jr_c, int p) {
        constexpr int JM = decltype(jm_c)::value;
        constexpr int JR = decltype(jr_c)::value;
        constexpr int NEW = JM % 3, OLD = (JM + 1) % 3, MID = (JM + 2) % 3;
        constexpr int B = JM & 1;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            ev[q][(JM + PF) % 6] = e[at(q, p + PF)];
            rv[q][(JM + 1) % 6] = r[at(q, p + 1)];
        }
#pragma unroll
        for (int l = 1; l <= K; ++l) {
            if (2 * l > JR) continue;
            const int zq = p - l;
            const bool zin = (unsigned)zq < (unsigned)n;
            float a[R], c[R], b[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                a[q] = l == 1 ? ev[q][(JM + 4) % 6] : w[l - 1][q][OLD];
                c[q] = l == 1 ? ev[q][(JM + 5) % 6] : w[l - 1][q][MID];
                b[q] = l == 1 ? ev[q][JM] : w[l - 1][q][NEW];
            }
            const float below = lds[B][l - 1][tyd][1][tx];  // last row of the thread row above
            const float above = lds[B][l - 1][tyu][0][tx];  // first row of the thread row below
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const float ym = q == 0 ? below : c[q - 1];
                const float yp = q == R - 1 ? above : c[q + 1];
                float sum = a[q];               // ( 0,  0, -1)
                sum += ym;                      // ( 0, -1,  0)
                sum += lane_prev(c[q]);         // (-1,  0,  0)
                sum += lane_next(c[q]);         // ( 1,  0,  0)
                sum += yp;                      // ( 0,  1,  0)
                sum += b[q];                    // ( 0,  0,  1)
                sum += rv[q][(JM + 6 - l) % 6]; // r at plane zq
                const float jv = DAMPED ? __builtin_fmaf(omega, __builtin_fmaf(sum, 1.0f / 6.0f, -c[q]), c[q])
                                        : sum * (1.0f / 6.0f);
                const float v = (inxy[q] && zin) ? jv : 0.0f;
                if (l < K) {
                    w[l][q][NEW] = v;
                } else if (writer[q]) {
                    e2[col[q] + (int64_t)(zq + 1) * sxy] = v;
                    if (partial) {
                        const double rr = (double)sum - 6.0 * (double)c[q];
                        acc += rr * rr;
                    }
                }
            }
        }
        if (JR >= 1) {
            lds[B ^ 1][0][ty][0][tx] = ev[0][JM];
            lds[B ^ 1][0][ty][1][tx] = ev[R - 1][JM];
        }
#pragma unroll
        for (int l = 1; l < K; ++l)
            if (2 * l + 1 <= JR) {
                lds[B ^ 1][l][ty][0][tx] = w[l][0][NEW];
                lds[B ^ 1][l][ty][1][tx] = w[l][R - 1][NEW];
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
    if (partial) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (tx == 0) red[ty] = acc;
        __syncthreads();
        if (tx == 0 && ty == 0) {
            double sum = 0.0;
            for (int k = 0; k < TY; ++k) sum += red[k];
            partial[tid] = sum;
        }
    }
}

// u += e over the interior and reset both correction buffers to zero
__global__ __launch_bounds__(kThreads) void poisson_update_f64(double *__restrict__ u, float *e, float *eother,
                                                               const Tile *__restrict__ tiles,
                                                               const ProblemDev *__restrict__ probs) {
    const Tile t = tiles[blockIdx.x];
    const ProblemDev P = probs[t.p];
    const int n = P.n;
    const int x = t.x0 + (int)(threadIdx.x % kTX);
    const int y = t.y0 + (int)(threadIdx.x / kTX);
    if (x >= n || y >= n) return;
    const int64_t sx = n + 2, sxy = sx * sx;
    const int z1 = min(t.z0 + kTZ, n);
    int64_t i = P.base + ((int64_t)(t.z0 + 1) * sx + (y + 1)) * sx + (x + 1);
    for (int z = t.z0; z < z1; ++z, i += sxy) {
        u[i] += (double)e[i];
        e[i] = 0.0f;
        eother[i] = 0.0f;
    }
}

// one plain FP64 Jacobi sweep on u (the final polishing sweeps):
// u'[c] = (u[z-1] + u[y-1] + u[x-1] + u[x+1] + u[y+1] + u[z+1] + f[c]) * (1/6),
// the six neighbours summed in that order, then f added, then one multiply
// (the oracle's plain FP64 Jacobi, lgdo_poisson_jacobi_f64)
__global__ __launch_bounds__(kThreads) void poisson_sweep_f64(const double *__restrict__ u, double *__restrict__ u2,
                                                              const float *__restrict__ f,
                                                              const Tile *__restrict__ tiles,
                                                              const ProblemDev *__restrict__ probs) {
    const Tile t = tiles[blockIdx.x];
    const ProblemDev P = probs[t.p];
    const int n = P.n;
    const int x = t.x0 + (int)(threadIdx.x % kTX);
    const int y = t.y0 + (int)(threadIdx.x / kTX);
    if (x >= n || y >= n) return;
    const int64_t sx = n + 2, sxy = sx * sx;
    const int z1 = min(t.z0 + kTZ, n);
    int64_t i = P.base + ((int64_t)(t.z0 + 1) * sx + (y + 1)) * sx + (x + 1);
    double below = u[i - sxy], center = u[i];
    for (int z = t.z0; z < z1; ++z, i += sxy) {
        const double above = u[i + sxy];
        double nb = below;
        nb += u[i - sx];
        nb += u[i - 1];
        nb += u[i + 1];
        nb += u[i + sx];
        nb += above;
        u2[i] = (nb + (double)f[i]) * (1.0 / 6.0);
        below = center;
        center = above;
    }
}

// norms[k] = sum of partial over the tiles [begin[k], end[k])  (fixed order)
__global__ __launch_bounds__(256) void poisson_reduce(const double *__restrict__ partial, const int *__restrict__ begin,
                                                      const int *__restrict__ end, double *__restrict__ norms) {
    __shared__ double lds[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int i = begin[k] + (int)threadIdx.x; i < end[k]; i += 256) s += partial[i];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) norms[k] = lds[0];
}

// faces into the halo of u (edges and corners are never read)
__global__ void poisson_faces(double *__restrict__ u, const ProblemDev *__restrict__ probs,
                              const double *__restrict__ faces) {
    const ProblemDev P = probs[blockIdx.y];
    const int n = P.n;
    const int64_t sx = n + 2;
    const double *fv = faces + 6 * blockIdx.y;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n * n; k += gridDim.x * blockDim.x) {
        const int a = k % n + 1, b = k / n + 1;
        u[P.base + ((int64_t)b * sx + a) * sx + 0] = fv[0];          // x = -1
        u[P.base + ((int64_t)b * sx + a) * sx + (n + 1)] = fv[1];    // x = n
        u[P.base + ((int64_t)b * sx + 0) * sx + a] = fv[2];          // y = -1
        u[P.base + ((int64_t)b * sx + (n + 1)) * sx + a] = fv[3];    // y = n
        u[P.base + ((int64_t)0 * sx + b) * sx + a] = fv[4];          // z = -1
        u[P.base + ((int64_t)(n + 1) * sx + b) * sx + a] = fv[5];    // z = n
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int64_t padded_elems(int n) { return (int64_t)(n + 2) * (n + 2) * (n + 2); }

int64_t tiles_of(int n) {
    return (int64_t)((n + kTX - 1) / kTX) * ((n + kTY - 1) / kTY) * ((n + kTZ - 1) / kTZ);
}

void append_tiles(std::vector<Tile> *out, int p, int n) {
    for (int z = 0; z < n; z += kTZ)
        for (int y = 0; y < n; y += kTY)
            for (int x = 0; x < n; x += kTX) out->push_back(Tile{p, x, y, z});
}

// temporal blocking shapes: variant -> (K, TY); tiles write (64-2K) x (TY-2K)
// columns and march z chunks of kTbZ planes (fewer on launches of few
// tiles, tb_zchunk)
struct TbShape {
    int k, ty, rows;  // sweeps per launch, work-group rows, tile rows per work-group row
};
constexpr int kTbZ = 64;
// the default: 4 sweeps per launch, three rows per thread in 48-row tiles
// (variant 6) for launches of at least kTallMinTiles of its tiles, else one
// row per thread (variant 3).  On MI355X the full 512-problem ensemble runs
// 3.81 ms per launch with three rows, 3.99 ms with two, 4.72 ms with one
// (50.0 / 52.2 / 62.0 s per solve), while 16 problems of 128^3 (576 two-row
// tiles, two rounds and a quarter of one work-group per CU) ran 186 us with
// two rows against 173 us with one: the one-row kernel holds two
// work-groups per CU, which hides more latency when a launch is only a few
// rounds long.
constexpr int kDefaultSweepVariant = 6;
constexpr int kSmallSweepVariant = 3;
constexpr int64_t kTallMinTiles = 2048;
constexpr int kSweepVariants = 6;
TbShape tb_shape(int variant) {
    switch (variant) {
    case 2: return {2, 16, 1};
    case 3: return {4, 16, 1};
    case 4: return {4, 16, 2};  // two rows per thread, 32-row tiles
    case 5: return {4, 8, 2};   // two rows per thread, 16-row tiles
    case 6: return {4, 16, 3};  // three rows per thread, 48-row tiles
    default: return {0, 0, 1};  // 1: plain sweeps
    }
}

int64_t tb_tiles_of(int n, TbShape sh, int zchunk = kTbZ) {
    const int ox = 64 - 2 * sh.k, oy = sh.ty * sh.rows - 2 * sh.k;
    return (int64_t)((n + ox - 1) / ox) * ((n + oy - 1) / oy) * ((n + zchunk - 1) / zchunk);
}

void append_tb_tiles(std::vector<Tile> *out, int p, int n, TbShape sh, int zchunk) {
    const int ox = 64 - 2 * sh.k, oy = sh.ty * sh.rows - 2 * sh.k;
    for (int z = 0; z < n; z += zchunk)
        for (int y = 0; y < n; y += oy)
            for (int x = 0; x < n; x += ox) out->push_back(Tile{p, x, y, z});
}

// z planes per blocked work-group for a launch over `set`: kTbZ, halved
// (down to 4) while the launch has fewer than kTbMinGroups tiles.  A tile
// is a z march of chunk + 2K plane steps, so a launch of small problems
// (16 coarse grids of 31^3: 64 tiles of 39 steps at kTbZ) takes as long as
// its longest march; 4-plane chunks give 512 tiles of 12 steps.  Blocked
// results do not depend on the chunk.  Halving stops once the count
// reaches kTbMinGroups, so a reduced list holds fewer than
// kTbSmallListTiles tiles (twice the count before the last halving plus
// one rounding tile per (problem, xy tile), itself under kTbMinGroups).
constexpr int64_t kTbMinGroups = 512;  // two one-row work-groups per CU
constexpr int64_t kTbSmallListTiles = 3 * kTbMinGroups;
int tb_zchunk(const lgdhip_poisson_problem *probs, const std::vector<int> &set, TbShape sh) {
    auto tiles = [&](int zc) {
        int64_t t = 0;
        for (int k : set) t += tb_tiles_of(probs[k].n, sh, zc);
        return t;
    };
    int zc = kTbZ;
    while (zc > 4 && tiles(zc) < kTbMinGroups) zc /= 2;
    return zc;
}

// the blocked sweep shape of a launch over the problems in `set`: the
// requested variant, or for 0 the default by launch size
TbShape sweep_shape(int requested, const lgdhip_poisson_problem *probs, const std::vector<int> &set) {
    if (requested >= 1 && requested <= kSweepVariants) return tb_shape(requested);
    const TbShape tall = tb_shape(kDefaultSweepVariant);
    int64_t tiles = 0;
    for (int k : set) tiles += tb_tiles_of(probs[k].n, tall);
    return tiles >= kTallMinTiles ? tall : tb_shape(kSmallSweepVariant);
}

int check_problems(const lgdhip_poisson_problem *probs, int count) {
    LGD_REQUIRE(probs && count > 0, "no problems");
    for (int k = 0; k < count; ++k) {
        if (probs[k].n < 1 || probs[k].n > 4096)
            return lgdhip::set_error(LGDHIP_ERR_INVALID_ARG, "problem edge length must be in [1, 4096]");
        if (probs[k].offset < 0) return lgdhip::set_error(LGDHIP_ERR_INVALID_ARG, "negative problem offset");
    }
    return 0;
}

// workspace carve-up, every piece 256-byte aligned
struct Workspace {
    float *e0, *e1, *r;
    double *partial, *norms;
    Tile *tiles;
    ProblemDev *probs;
    int *begin, *end;
    double *faces;
    Tile *tb_tiles;  // temporal-blocking sweep tiles of the active set
    int *tb_begin, *tb_end;
    size_t bytes;
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

Workspace carve(void *base, int64_t pool, int64_t ntiles, int64_t ntb, int count) {
    Workspace w;
    char *p = (char *)base;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *q = p ? p + off : nullptr;
        off += align256(bytes);
        return q;
    };
    w.e0 = (float *)take(pool * 4);
    w.e1 = (float *)take(pool * 4);
    w.r = (float *)take(pool * 4);
    w.partial = (double *)take(std::max(ntiles, ntb) * 8);
    w.norms = (double *)take((size_t)count * 8);
    w.tiles = (Tile *)take(ntiles * sizeof(Tile));
    w.probs = (ProblemDev *)take((size_t)count * sizeof(ProblemDev));
    w.begin = (int *)take((size_t)count * 4);
    w.end = (int *)take((size_t)count * 4);
    w.faces = (double *)take((size_t)count * 6 * 8);
    w.tb_tiles = (Tile *)take(ntb * sizeof(Tile));
    w.tb_begin = (int *)take((size_t)count * 4);
    w.tb_end = (int *)take((size_t)count * 4);
    w.bytes = off;
    return w;
}

int64_t pool_extent(const lgdhip_poisson_problem *probs, int count) {
    int64_t e = 0;
    for (int k = 0; k < count; ++k) e = std::max(e, probs[k].offset + padded_elems(probs[k].n));
    return e;
}

int64_t total_tiles(const lgdhip_poisson_problem *probs, int count) {
    int64_t t = 0;
    for (int k = 0; k < count; ++k) t += tiles_of(probs[k].n);
    return t;
}

// tile capacity for every temporal-blocking shape: kTbZ chunks, or a
// list of reduced chunks (under kTbSmallListTiles, tb_zchunk)
int64_t total_tb_tiles(const lgdhip_poisson_problem *probs, int count) {
    int64_t t = 0;
    for (int k = 0; k < count; ++k) {
        int64_t t_max = 0;  // tile lists of every blocked variant fit
        for (int v = 2; v <= kSweepVariants; ++v) t_max = std::max(t_max, tb_tiles_of(probs[k].n, tb_shape(v)));
        t += t_max;
    }
    return std::max(t, kTbSmallListTiles);
}

Workspace carve_for(void *base, const lgdhip_poisson_problem *probs, int count) {
    return carve(base, pool_extent(probs, count), total_tiles(probs, count), total_tb_tiles(probs, count), count);
}

// Uploads the temporal-blocking tile list of the problems in `set` (z
// chunks of tb_zchunk planes, returned in *zchunk for the launches);
// tb_begin/tb_end index the tiles of set[k].  Returns the tile count.
int upload_tb_set(const Workspace &w, const lgdhip_poisson_problem *probs, const std::vector<int> &set, TbShape sh,
                  int64_t *ntiles, int *zchunk, hipStream_t s) {
    std::vector<Tile> tiles;
    std::vector<int> b(set.size()), e(set.size());
    *zchunk = tb_zchunk(probs, set, sh);
    for (size_t k = 0; k < set.size(); ++k) {
        b[k] = (int)tiles.size();
        append_tb_tiles(&tiles, set[k], probs[set[k]].n, sh, *zchunk);
        e[k] = (int)tiles.size();
    }
    *ntiles = (int64_t)tiles.size();
    if (tiles.empty()) return 0;
    LGD_HIP(hipMemcpyAsync(w.tb_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, s));
    LGD_HIP(hipMemcpyAsync(w.tb_begin, b.data(), b.size() * 4, hipMemcpyHostToDevice, s));
    LGD_HIP(hipMemcpyAsync(w.tb_end, e.data(), e.size() * 4, hipMemcpyHostToDevice, s));
    LGD_HIP(hipStreamSynchronize(s));
    return 0;
}

template <bool DAMPED>
void launch_tb_kind(TbShape sh, const float *a, float *b, const float *r, const Workspace &w, int64_t ntiles,
                    int zchunk, double *partial, float omega, hipStream_t s) {
    const unsigned nt = (unsigned)ntiles, per = lgdhip::xcd_per(nt), groups = lgdhip::xcd_groups(nt);
    if (sh.rows == 3)
        poisson_sweep_tb2<4, 16, 3, DAMPED><<<groups, dim3(64, 16), 0, s>>>(a, b, r, w.tb_tiles, w.probs, partial,
                                                                            zchunk, nt, per, omega);
    else if (sh.rows == 2 && sh.ty == 16)
        poisson_sweep_tb2<4, 16, 2, DAMPED><<<groups, dim3(64, 16), 0, s>>>(a, b, r, w.tb_tiles, w.probs, partial,
                                                                            zchunk, nt, per, omega);
    else if (sh.rows == 2)
        poisson_sweep_tb2<4, 8, 2, DAMPED><<<groups, dim3(64, 8), 0, s>>>(a, b, r, w.tb_tiles, w.probs, partial, zchunk,
                                                                          nt, per, omega);
    else if (sh.k == 2)
        poisson_sweep_tb<2, 16, DAMPED><<<groups, dim3(64, 16), 0, s>>>(a, b, r, w.tb_tiles, w.probs, partial, zchunk, nt,
                                                                        per, omega);
    else
        poisson_sweep_tb<4, 16, DAMPED><<<groups, dim3(64, 16), 0, s>>>(a, b, r, w.tb_tiles, w.probs, partial, zchunk, nt,
                                                                        per, omega);
}

// one launch of K blocked sweeps a -> b over the uploaded tile list (z
// chunks of `zchunk` planes, as uploaded); omega > 0 runs weighted sweeps
// e' = e + omega (J(e) - e)
void launch_tb(TbShape sh, const float *a, float *b, const float *r, const Workspace &w, int64_t ntiles,
               int zchunk, double *partial, hipStream_t s, float omega = 0.0f) {
    if (ntiles <= 0) return;
    if (omega > 0.0f)
        launch_tb_kind<true>(sh, a, b, r, w, ntiles, zchunk, partial, omega, s);
    else
        launch_tb_kind<false>(sh, a, b, r, w, ntiles, zchunk, partial, 0.0f, s);
}

// Uploads problem descriptors and the tile list of the problems in `set`;
// begin/end index the tiles of set[k].
int upload_set(const Workspace &w, const lgdhip_poisson_problem *probs, const std::vector<int> &set,
               std::vector<Tile> *tiles, hipStream_t s) {
    tiles->clear();
    std::vector<int> b(set.size()), e(set.size());
    for (size_t k = 0; k < set.size(); ++k) {
        b[k] = (int)tiles->size();
        append_tiles(tiles, set[k], probs[set[k]].n);
        e[k] = (int)tiles->size();
    }
    if (tiles->empty()) return 0;
    LGD_HIP(hipMemcpyAsync(w.tiles, tiles->data(), tiles->size() * sizeof(Tile), hipMemcpyHostToDevice, s));
    LGD_HIP(hipMemcpyAsync(w.begin, b.data(), b.size() * 4, hipMemcpyHostToDevice, s));
    LGD_HIP(hipMemcpyAsync(w.end, e.data(), e.size() * 4, hipMemcpyHostToDevice, s));
    LGD_HIP(hipStreamSynchronize(s));
    return 0;
}

int residual_norms(const double *u, const float *f, const Workspace &w, const std::vector<Tile> &tiles, int nset,
                   std::vector<double> *norms, hipStream_t s) {
    if (tiles.empty()) return 0;
    poisson_residual_f64<<<(unsigned)tiles.size(), kThreads, 0, s>>>(u, f, w.r, w.tiles, w.probs, w.partial);
    LGD_LAUNCHED("poisson_residual_f64");
    poisson_reduce<<<(unsigned)nset, 256, 0, s>>>(w.partial, w.begin, w.end, w.norms);
    LGD_LAUNCHED("poisson_reduce");
    norms->resize(nset);
    LGD_HIP(hipMemcpyAsync(norms->data(), w.norms, (size_t)nset * 8, hipMemcpyDeviceToHost, s));
    LGD_HIP(hipStreamSynchronize(s));
    for (double &v : *norms) v = std::sqrt(v);
    return 0;
}

} // namespace

// `sweeps` plain (omega == 0) or weighted sweeps ea <-> eb over every
// problem: blocked launches of K sweeps, then single sweeps for the rest.
// Uses the descriptor and tile-list pieces of `w` (probs, tiles, begin, end,
// tb_*), which must hold every problem's lists.
int run_sweeps_in(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs, int count, int sweeps,
                  float omega, const lgdhip_tile *tile, const Workspace &w, int *result_in_b, hipStream_t s) {
    int rc = 0;
    std::vector<ProblemDev> pd(count);
    for (int k = 0; k < count; ++k) pd[k] = ProblemDev{probs[k].n, 0, probs[k].offset};
    LGD_HIP(hipMemcpyAsync(w.probs, pd.data(), pd.size() * sizeof(ProblemDev), hipMemcpyHostToDevice, s));
    std::vector<int> all(count);
    for (int k = 0; k < count; ++k) all[k] = k;
    std::vector<Tile> tiles;
    rc = upload_set(w, probs, all, &tiles, s);
    if (rc) return rc;
    const TbShape sh = sweep_shape(tile ? tile->variant : 0, probs, all);
    float *cur = ea, *nxt = eb;
    int t = 0;
    if (sh.k > 0 && sweeps >= sh.k) {
        int64_t ntb = 0;
        int zc = kTbZ;
        rc = upload_tb_set(w, probs, all, sh, &ntb, &zc, s);
        if (rc) return rc;
        for (; t + sh.k <= sweeps; t += sh.k) {
            launch_tb(sh, cur, nxt, r, w, ntb, zc, nullptr, s, omega);
            LGD_LAUNCHED("poisson_sweep_tb");
            std::swap(cur, nxt);
        }
    }
    for (; t < sweeps; ++t) {
        if (omega > 0.0f)
            poisson_sweep_f32<true>
                <<<(unsigned)tiles.size(), kThreads, 0, s>>>(cur, nxt, r, w.tiles, w.probs, nullptr, omega);
        else
            poisson_sweep_f32<false>
                <<<(unsigned)tiles.size(), kThreads, 0, s>>>(cur, nxt, r, w.tiles, w.probs, nullptr, 1.0f);
        LGD_LAUNCHED("poisson_sweep_f32");
        std::swap(cur, nxt);
    }
    if (result_in_b) *result_in_b = cur == eb ? 1 : 0;
    return 0;
}

int run_sweeps(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs, int count, int sweeps,
               float omega, const lgdhip_tile *tile, void *workspace, size_t workspace_bytes, int *result_in_b,
               lgdhip_stream_t stream) {
    int rc = check_problems(probs, count);
    if (rc) return rc;
    LGD_REQUIRE(ea && eb && ea != eb && r && workspace && sweeps >= 0, "bad arguments");
    const Workspace w = carve_for(workspace, probs, count);
    if (workspace_bytes < w.bytes) return lgdhip::set_error(LGDHIP_ERR_WORKSPACE, "poisson workspace too small");
    return run_sweeps_in(ea, eb, r, probs, count, sweeps, omega, tile, w, result_in_b, lgdhip::as_stream(stream));
}

// ---------------------------------------------------------------------------
// the no-allocation sweep entry point
// ---------------------------------------------------------------------------
// Every device buffer of lgdhip_poisson_sweeps_f32_noalloc beyond the
// caller's e, r pools is the problem descriptors and tile lists of one call;
// they live in this static array.  Each device holds its own copy for the
// life of the process; its address on a device comes from
// hipGetSymbolAddress (the array's name is a host-side stand-in, not a
// device address).  Calls on one device share it: overlapping calls there
// race.  Nothing in it needs to start at zero (every piece a call reads is
// uploaded by that call first).
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices];

// The device address of `scratch` on the current device, resolved once per
// device on its first use.
int scratch_on_current_device(void **out) {
    int dev = -1;
    LGD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return lgdhip::set_error(LGDHIP_ERR_INVALID_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return lgdhip::hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// The descriptor and tile-list pieces run_sweeps_in uses, carved from
// `base` (nullptr: sizes only) for `ntiles` single-sweep and `ntb` blocked
// tiles of `count` problems; the pool pieces stay null.
Workspace carve_lists(void *base, int64_t ntiles, int64_t ntb, int count) {
    Workspace w{};
    char *p = (char *)base;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *q = p ? p + off : nullptr;
        off += align256(bytes);
        return q;
    };
    w.tiles = (Tile *)take(ntiles * sizeof(Tile));
    w.probs = (ProblemDev *)take((size_t)count * sizeof(ProblemDev));
    w.begin = (int *)take((size_t)count * 4);
    w.end = (int *)take((size_t)count * 4);
    w.tb_tiles = (Tile *)take(ntb * sizeof(Tile));
    w.tb_begin = (int *)take((size_t)count * 4);
    w.tb_end = (int *)take((size_t)count * 4);
    w.bytes = off;
    return w;
}
} // namespace

extern "C" {

int64_t lgdhip_poisson_layout(lgdhip_poisson_problem *probs, int count) {
    if (!probs || count <= 0) return 0;
    int64_t off = 0;
    for (int k = 0; k < count; ++k) {
        probs[k].offset = off;
        // keep every problem block 256-byte aligned in the FP32 pools
        off += (padded_elems(probs[k].n) + 63) / 64 * 64;
    }
    return off;
}

size_t lgdhip_poisson_workspace_bytes(const lgdhip_poisson_problem *probs, int count) {
    if (check_problems(probs, count)) return 0;
    return carve_for(nullptr, probs, count).bytes;
}

int lgdhip_poisson_set_faces(double *u, const lgdhip_poisson_problem *probs, int count, void *workspace,
                             size_t workspace_bytes, lgdhip_stream_t stream) {
    int rc = check_problems(probs, count);
    if (rc) return rc;
    LGD_REQUIRE(u && workspace, "NULL pointer");
    const Workspace w = carve_for(workspace, probs, count);
    if (workspace_bytes < w.bytes) return lgdhip::set_error(LGDHIP_ERR_WORKSPACE, "poisson workspace too small");
    hipStream_t s = lgdhip::as_stream(stream);
    std::vector<ProblemDev> pd(count);
    std::vector<double> faces(6 * (size_t)count);
    int nmax = 1;
    for (int k = 0; k < count; ++k) {
        pd[k] = ProblemDev{probs[k].n, 0, probs[k].offset};
        std::memcpy(&faces[6 * (size_t)k], probs[k].face, 6 * sizeof(double));
        nmax = std::max(nmax, probs[k].n);
    }
    LGD_HIP(hipMemcpyAsync(w.probs, pd.data(), pd.size() * sizeof(ProblemDev), hipMemcpyHostToDevice, s));
    LGD_HIP(hipMemcpyAsync(w.faces, faces.data(), faces.size() * 8, hipMemcpyHostToDevice, s));
    LGD_REQUIRE(count <= 65535, "at most 65535 problems");
    dim3 grid(lgdhip::ceil_div((int64_t)nmax * nmax, 256), (unsigned)count);
    poisson_faces<<<grid, 256, 0, s>>>(u, w.probs, w.faces);
    LGD_LAUNCHED("poisson_faces");
    LGD_HIP(hipStreamSynchronize(s));
    return 0;
}

int lgdhip_poisson_sweeps_f32(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs, int count,
                              int sweeps, const lgdhip_tile *tile, void *workspace, size_t workspace_bytes,
                              int *result_in_b, lgdhip_stream_t stream) {
    return run_sweeps(ea, eb, r, probs, count, sweeps, 0.0f, tile, workspace, workspace_bytes, result_in_b, stream);
}

int lgdhip_poisson_damped_sweeps_f32(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs,
                                     int count, int sweeps, float omega, const lgdhip_tile *tile, void *workspace,
                                     size_t workspace_bytes, int *result_in_b, lgdhip_stream_t stream) {
    LGD_REQUIRE(omega > 0.0f && omega <= 1.0f, "omega must be in (0, 1]");
    return run_sweeps(ea, eb, r, probs, count, sweeps, omega, tile, workspace, workspace_bytes, result_in_b, stream);
}

int lgdhip_poisson_sweeps_f32_noalloc(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs,
                                      int count, int sweeps, const lgdhip_tile *tile, int *result_in_b,
                                      lgdhip_stream_t stream) {
    int rc = check_problems(probs, count);
    if (rc) return rc;
    LGD_REQUIRE(ea && eb && ea != eb && r && sweeps >= 0, "bad arguments");
    // the lists this call uploads: every problem's single-sweep tiles and the
    // blocked tiles of the shape the call runs
    std::vector<int> all(count);
    for (int k = 0; k < count; ++k) all[k] = k;
    const TbShape sh = sweep_shape(tile ? tile->variant : 0, probs, all);
    int64_t ntb = 0;
    if (sh.k > 0)
        for (int k = 0; k < count; ++k) ntb += tb_tiles_of(probs[k].n, sh, tb_zchunk(probs, all, sh));
    const size_t need = carve_lists(nullptr, total_tiles(probs, count), ntb, count).bytes;
    if (need > kNoAllocScratchBytes)
        return lgdhip::set_error(LGDHIP_ERR_WORKSPACE,
                                 "problem set needs " + std::to_string(need) + " bytes of tile lists; the static "
                                 "scratch holds " + std::to_string(kNoAllocScratchBytes));
    void *base = nullptr;
    rc = scratch_on_current_device(&base);
    if (rc) return rc;
    const Workspace w = carve_lists(base, total_tiles(probs, count), ntb, count);
    return run_sweeps_in(ea, eb, r, probs, count, sweeps, 0.0f, tile, w, result_in_b, lgdhip::as_stream(stream));
}

int lgdhip_poisson_sweeps_f64(double *ua, double *ub, const float *f, const lgdhip_poisson_problem *probs,
                              int count, int sweeps, void *workspace, size_t workspace_bytes, int *result_in_b,
                              lgdhip_stream_t stream) {
    int rc = check_problems(probs, count);
    if (rc) return rc;
    LGD_REQUIRE(ua && ub && ua != ub && f && workspace && sweeps >= 0, "bad arguments");
    const Workspace w = carve_for(workspace, probs, count);
    if (workspace_bytes < w.bytes) return lgdhip::set_error(LGDHIP_ERR_WORKSPACE, "poisson workspace too small");
    hipStream_t s = lgdhip::as_stream(stream);
    std::vector<ProblemDev> pd(count);
    for (int k = 0; k < count; ++k) pd[k] = ProblemDev{probs[k].n, 0, probs[k].offset};
    LGD_HIP(hipMemcpyAsync(w.probs, pd.data(), pd.size() * sizeof(ProblemDev), hipMemcpyHostToDevice, s));
    std::vector<int> all(count);
    for (int k = 0; k < count; ++k) all[k] = k;
    std::vector<Tile> tiles;
    rc = upload_set(w, probs, all, &tiles, s);
    if (rc) return rc;
    double *cur = ua, *nxt = ub;
    for (int t = 0; t < sweeps; ++t) {
        poisson_sweep_f64<<<(unsigned)tiles.size(), kThreads, 0, s>>>(cur, nxt, f, w.tiles, w.probs);
        LGD_LAUNCHED("poisson_sweep_f64");
        std::swap(cur, nxt);
    }
    if (result_in_b) *result_in_b = cur == ub ? 1 : 0;
    return 0;
}

int lgdhip_poisson_residual_norms(const double *u, const float *f, const lgdhip_poisson_problem *probs, int count,
                                  void *workspace, size_t workspace_bytes, double *norms,
                                  lgdhip_stream_t stream) {
    int rc = check_problems(probs, count);
    if (rc) return rc;
    LGD_REQUIRE(u && f && workspace && norms, "NULL pointer");
    const Workspace w = carve_for(workspace, probs, count);
    if (workspace_bytes < w.bytes) return lgdhip::set_error(LGDHIP_ERR_WORKSPACE, "poisson workspace too small");
    hipStream_t s = lgdhip::as_stream(stream);
    std::vector<ProblemDev> pd(count);
    for (int k = 0; k < count; ++k) pd[k] = ProblemDev{probs[k].n, 0, probs[k].offset};
    LGD_HIP(hipMemcpyAsync(w.probs, pd.data(), pd.size() * sizeof(ProblemDev), hipMemcpyHostToDevice, s));
    std::vector<int> all(count);
    for (int k = 0; k < count; ++k) all[k] = k;
    std::vector<Tile> tiles;
    rc = upload_set(w, probs, all, &tiles, s);
    if (rc) return rc;
    std::vector<double> nv;
    rc = residual_norms(u, f, w, tiles, count, &nv, s);
    if (rc) return rc;
    std::memcpy(norms, nv.data(), nv.size() * 8);
    return 0;
}

int lgdhip_poisson_sweep_plan(const lgdhip_poisson_problem *probs, int count, int requested, int *variant) {
    int rc = check_problems(probs, count);
    if (rc) return rc;
    LGD_REQUIRE(variant, "NULL output pointer");
    if (requested >= 1 && requested <= kSweepVariants) {
        *variant = requested;
        return 0;
    }
    std::vector<int> all(count);
    for (int k = 0; k < count; ++k) all[k] = k;
    const TbShape sh = sweep_shape(0, probs, all);
    *variant = sh.rows > 1 ? kDefaultSweepVariant : kSmallSweepVariant;
    return 0;
}

int lgdhip_poisson_batch_solve(double *u, const float *f, const lgdhip_poisson_problem *probs, int count,
                               const lgdhip_poisson_options *opt, void *workspace, size_t workspace_bytes,
                               lgdhip_poisson_result *results, lgdhip_stream_t stream) {
    int rc = check_problems(probs, count);
    if (rc) return rc;
    LGD_REQUIRE(u && f && workspace && results && opt, "NULL pointer");
    LGD_REQUIRE(opt->rel_tol > 0 && opt->chunk >= 1 && opt->max_sweeps >= 1, "bad solver options");
    LGD_REQUIRE(opt->smooth_sweeps >= 0 && (opt->smooth_sweeps == 0 || (opt->smooth_omega > 0.0f &&
                                                                        opt->smooth_omega <= 1.0f)),
                "smooth_omega must be in (0, 1]");
    const Workspace w = carve_for(workspace, probs, count);
    if (workspace_bytes < w.bytes) return lgdhip::set_error(LGDHIP_ERR_WORKSPACE, "poisson workspace too small");
    hipStream_t s = lgdhip::as_stream(stream);
    const bool graphs = opt->use_graph != 0 && s != nullptr;
    const double theta = opt->inner_reduction > 0 ? opt->inner_reduction : 1e-2;
    // the blocked shape is chosen per active set (sweep_shape); the default's
    // two candidates both run K = 4, so the chunk length below holds for any
    TbShape sh = sweep_shape(opt->sweep_variant, probs, std::vector<int>());
    // sweeps per chunk: even (plain) or a multiple of 2K (blocked), so every
    // chunk leaves the correction in the same buffer
    const int unit = sh.k > 0 ? 2 * sh.k : 2;
    const int C = std::max(unit, (opt->chunk + unit - 1) / unit * unit);

    std::vector<ProblemDev> pd(count);
    for (int k = 0; k < count; ++k) pd[k] = ProblemDev{probs[k].n, 0, probs[k].offset};
    LGD_HIP(hipMemcpyAsync(w.probs, pd.data(), pd.size() * sizeof(ProblemDev), hipMemcpyHostToDevice, s));

    // zero the correction pools (a kernel fill, never a memset node)
    const int64_t pool = pool_extent(probs, count);
    rc = lgdhip_fill_u8(w.e0, 0, (size_t)pool * 4, stream);
    if (rc) return rc;
    rc = lgdhip_fill_u8(w.e1, 0, (size_t)pool * 4, stream);
    if (rc) return rc;

    // FP64 residual of the caller's initial guess.  The tolerance is relative
    // to opt->reference_norms[k] when given (e.g. the zero-guess residual when
    // u holds a predicted guess), else to this initial residual.
    std::vector<int> active(count);
    for (int k = 0; k < count; ++k) active[k] = k;
    std::vector<Tile> tiles;
    rc = upload_set(w, probs, active, &tiles, s);
    if (rc) return rc;
    std::vector<double> rn;
    rc = residual_norms(u, f, w, tiles, count, &rn, s);
    if (rc) return rc;

    struct State {
        double ref, outer, inner;
        int64_t sweeps, inner_sweeps;
        int outer_iters;
        bool done;
    };
    std::vector<State> st(count);
    for (int k = 0; k < count; ++k) {
        const double ref = (opt->reference_norms && opt->reference_norms[k] > 0) ? opt->reference_norms[k] : rn[k];
        st[k] = State{ref, rn[k], rn[k], 0, 0, 0, rn[k] <= opt->rel_tol * ref};
    }

    float *ecur = w.e0, *enxt = w.e1;  // every problem's correction lives in ecur between chunks

    // smoothing phase: weighted Jacobi on the correction of every problem
    // not yet converged (an even count leaves the correction in ecur); the
    // main loop then continues plain Jacobi from that correction
    const int smooth = opt->smooth_sweeps + (opt->smooth_sweeps & 1);
    if (smooth > 0) {
        active.clear();
        for (int k = 0; k < count; ++k)
            if (!st[k].done) active.push_back(k);
        if (!active.empty()) {
            rc = upload_set(w, probs, active, &tiles, s);
            if (rc) return rc;
            // blocked weighted sweeps in pairs of launches (each launch swaps
            // the buffers once), single weighted sweeps for the rest
            const TbShape ssh = sweep_shape(opt->sweep_variant, probs, active);
            int t = 0;
            if (ssh.k > 0 && smooth >= 2 * ssh.k) {
                int64_t ntb = 0;
                int zc = kTbZ;
                rc = upload_tb_set(w, probs, active, ssh, &ntb, &zc, s);
                if (rc) return rc;
                for (; t + 2 * ssh.k <= smooth; t += 2 * ssh.k) {
                    for (int half = 0; half < 2; ++half) {
                        launch_tb(ssh, ecur, enxt, w.r, w, ntb, zc, nullptr, s, opt->smooth_omega);
                        LGD_LAUNCHED("poisson_sweep_tb");
                        std::swap(ecur, enxt);
                    }
                }
            }
            for (; t < smooth; ++t) {
                poisson_sweep_f32<true><<<(unsigned)tiles.size(), kThreads, 0, s>>>(
                    ecur, enxt, w.r, w.tiles, w.probs, nullptr, opt->smooth_omega);
                LGD_LAUNCHED("poisson_sweep_f32");
                std::swap(ecur, enxt);
            }
            for (int k : active) {
                st[k].sweeps += smooth;
                st[k].inner_sweeps += smooth;
            }
        }
    }
    // the chunk graph, destroyed on every return path
    struct GraphExec {
        hipGraphExec_t h = nullptr;
        void reset() {
            if (h) hipGraphExecDestroy(h);
            h = nullptr;
        }
        ~GraphExec() { reset(); }
    } graph;
    bool graph_valid = false;
    while (true) {
        active.clear();
        for (int k = 0; k < count; ++k)
            if (!st[k].done) active.push_back(k);
        if (active.empty()) break;
        rc = upload_set(w, probs, active, &tiles, s);
        if (rc) return rc;
        int64_t ntb = 0;
        int zc = kTbZ;
        sh = sweep_shape(opt->sweep_variant, probs, active);
        if (sh.k > 0) {
            rc = upload_tb_set(w, probs, active, sh, &ntb, &zc, s);
            if (rc) return rc;
        }
        graph_valid = false;

        bool set_changed = false;
        while (!set_changed) {
            // one chunk of C sweeps whose last sweep also reports the inner
            // residual of its input iterate; the chunk leaves the correction
            // in ecur, so the captured graph stays valid from chunk to chunk
            auto enqueue = [&](float *a, float *b) -> int {
                if (sh.k > 0) {
                    for (int t = 0; t < C; t += sh.k) {
                        launch_tb(sh, a, b, w.r, w, ntb, zc, t + sh.k == C ? w.partial : nullptr, s);
                        std::swap(a, b);
                    }
                    poisson_reduce<<<(unsigned)active.size(), 256, 0, s>>>(w.partial, w.tb_begin, w.tb_end,
                                                                           w.norms);
                    return 0;
                }
                for (int t = 0; t < C; ++t) {
                    poisson_sweep_f32<false><<<(unsigned)tiles.size(), kThreads, 0, s>>>(
                        a, b, w.r, w.tiles, w.probs, t == C - 1 ? w.partial : nullptr, 1.0f);
                    std::swap(a, b);
                }
                poisson_reduce<<<(unsigned)active.size(), 256, 0, s>>>(w.partial, w.begin, w.end, w.norms);
                return 0;
            };
            if (graphs) {
                if (!graph_valid) {
                    graph.reset();
                    LGD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    enqueue(ecur, enxt);
                    hipGraph_t g = nullptr;
                    LGD_HIP(hipStreamEndCapture(s, &g));
                    hipError_t e = hipGraphInstantiate(&graph.h, g, nullptr, nullptr, 0);
                    hipGraphDestroy(g);
                    if (e != hipSuccess) {
                        graph.h = nullptr;
                        return lgdhip::hip_fail(e, "hipGraphInstantiate");
                    }
                    graph_valid = true;
                }
                LGD_HIP(hipGraphLaunch(graph.h, s));
            } else {
                enqueue(ecur, enxt);
                LGD_LAUNCHED("poisson_sweep_f32");
            }
            std::vector<double> inner(active.size());
            LGD_HIP(hipMemcpyAsync(inner.data(), w.norms, inner.size() * 8, hipMemcpyDeviceToHost, s));
            LGD_HIP(hipStreamSynchronize(s));

            // problems whose correction has converged far enough get an
            // FP64 update u += e and a fresh FP64 residual
            std::vector<int> refresh;
            for (size_t a = 0; a < active.size(); ++a) {
                State &q = st[active[a]];
                q.sweeps += C;
                q.inner_sweeps += C;
                q.inner = std::sqrt(inner[a]);
                const bool inner_ok = q.inner <= theta * q.outer || q.inner <= 0.5 * opt->rel_tol * q.ref;
                if (inner_ok || q.sweeps >= opt->max_sweeps) refresh.push_back(active[a]);
            }
            if (refresh.empty()) continue;
            std::vector<Tile> rt;
            rc = upload_set(w, probs, refresh, &rt, s);
            if (rc) return rc;
            poisson_update_f64<<<(unsigned)rt.size(), kThreads, 0, s>>>(u, ecur, enxt, w.tiles, w.probs);
            LGD_LAUNCHED("poisson_update_f64");
            rc = residual_norms(u, f, w, rt, (int)refresh.size(), &rn, s);
            if (rc) return rc;
            for (size_t a = 0; a < refresh.size(); ++a) {
                State &q = st[refresh[a]];
                q.outer = rn[a];
                q.outer_iters += 1;
                q.inner_sweeps = 0;
                if (q.outer <= opt->rel_tol * q.ref || q.sweeps >= opt->max_sweeps) {
                    q.done = true;
                    set_changed = true;
                }
            }
            // the refreshed problems restart their correction from zero; the
            // tile list on the device was replaced, so re-upload the active set
            rc = upload_set(w, probs, active, &tiles, s);
            if (rc) return rc;
            graph_valid = false;
        }
    }
    graph.reset();
    // final FP64 polishing: plain FP64 Jacobi sweeps on u itself for every
    // problem (an even count, so u holds the result), the second buffer being
    // the correction pools' storage (no longer needed), then fresh FP64
    // residuals
    const int polish = opt->final_fp64_sweeps + (opt->final_fp64_sweeps & 1);
    if (polish > 0) {
        double *u2 = reinterpret_cast<double *>(w.e0);  // e0 and e1 together hold pool doubles
        LGD_HIP(hipMemcpyAsync(u2, u, (size_t)pool * sizeof(double), hipMemcpyDeviceToDevice, s));
        active.resize(count);
        for (int k = 0; k < count; ++k) active[k] = k;
        rc = upload_set(w, probs, active, &tiles, s);
        if (rc) return rc;
        double *cur = u, *nxt = u2;
        for (int t = 0; t < polish; ++t) {
            poisson_sweep_f64<<<(unsigned)tiles.size(), kThreads, 0, s>>>(cur, nxt, f, w.tiles, w.probs);
            LGD_LAUNCHED("poisson_sweep_f64");
            std::swap(cur, nxt);
        }
        rc = residual_norms(u, f, w, tiles, count, &rn, s);
        if (rc) return rc;
        for (int k = 0; k < count; ++k) {
            st[k].outer = rn[k];
            st[k].sweeps += polish;
        }
    }
    for (int k = 0; k < count; ++k) {
        results[k].sweeps = st[k].sweeps;
        results[k].outer_iterations = st[k].outer_iters;
        results[k].converged = st[k].outer <= opt->rel_tol * st[k].ref ? 1 : 0;
        results[k].reference_norm = st[k].ref;
        results[k].residual_norm = st[k].outer;
    }
    return 0;
}

} // extern "C"
