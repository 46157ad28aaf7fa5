// Conway's Game of Life (B3/S23, Moore<2,1>, Cube topology with a constant
// edge state) on gfx950.
//
// Cell semantics follow ConwayCell::update (src/examples/gameoflife/main.cpp:36-59):
// count the eight Moore neighbours, a live cell survives with 2 or 3 live
// neighbours, a dead cell is born with exactly 3.  Equivalently, with the
// 3x3 box sum s9 (centre included): alive' = (s9 == 3) | (alive & s9 == 4),
// which is what both kernels below evaluate.
//
// Two device representations:
//  * the user-visible uint8 grid (one byte per cell, 0/1) in the padded SoA
//    layout, swept by a SWAR kernel (four cells per 32-bit lane word);
//  * a bit-packed state (32 cells per word) swept by a register-resident
//    temporal-blocking kernel: every lane owns one word column of R rows,
//    exchanges its left/right words with the neighbouring lanes and runs up
//    to 32 generations per launch without touching memory.  The uint8 grid
//    is packed once before and unpacked once after a run.
#include "lgdhip_internal.h"

namespace {

// ---------------------------------------------------------------------------
// byte-per-cell SWAR kernel
// ---------------------------------------------------------------------------

// 0x80 in every byte of v that is zero; exact for bytes below 0x80.
__device__ __forceinline__ uint32_t zero_bytes(uint32_t v) {
    return ~(((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;
}

// Per-byte sum of the west, centre and east cells of four consecutive cells.
__device__ __forceinline__ uint32_t hsum3(const uint8_t *__restrict__ p) {
    const uint32_t w = *reinterpret_cast<const uint32_t *>(p);
    const uint32_t l = p[-1];
    const uint32_t r = p[4];
    return w + ((w << 8) | l) + ((w >> 8) | (r << 24));
}

template <int BY>
__global__ __launch_bounds__(64 * BY) void life_u8_swar(const uint8_t *__restrict__ src,
                                                        uint8_t *__restrict__ dst, int nx, int ny,
                                                        int64_t sx, int64_t origin) {
    const int x0 = 4 * (blockIdx.x * 64 + threadIdx.x);
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x0 >= nx || y >= ny) return;
    const int64_t i = origin + (int64_t)y * sx + x0;
    const uint32_t t = hsum3(src + i - sx) + hsum3(src + i) + hsum3(src + i + sx);
    const uint32_t c = *reinterpret_cast<const uint32_t *>(src + i);
    const uint32_t born = zero_bytes(t ^ 0x03030303u);
    const uint32_t stay = zero_bytes(t ^ 0x04040404u) & (c << 7);
    uint32_t res = (born | stay) >> 7;
    if (x0 + 4 > nx) {
        // last partial word: the bytes past nx are halo cells, keep them
        const uint32_t mask = (1u << (8 * (nx - x0))) - 1u;
        const uint32_t old = *reinterpret_cast<const uint32_t *>(dst + i);
        res = (res & mask) | (old & ~mask);
    }
    *reinterpret_cast<uint32_t *>(dst + i) = res;
}

// ---------------------------------------------------------------------------
// packing
// ---------------------------------------------------------------------------
__global__ void life_pack(const uint8_t *__restrict__ grid, uint32_t *__restrict__ packed, int nx, int ny,
                          int nw, int64_t sx, int64_t origin) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (w >= nw || y >= ny) return;
    const uint8_t *row = grid + origin + (int64_t)y * sx;
    const int x0 = 32 * w;
    uint32_t bits = 0;
    if (x0 + 32 <= nx) {
        const uint4 lo = *reinterpret_cast<const uint4 *>(row + x0);
        const uint4 hi = *reinterpret_cast<const uint4 *>(row + x0 + 16);
        const uint32_t q[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            // 0x80 in every non-zero byte, then gather the four flags
            const uint32_t v = q[k];
            const uint32_t nz = (((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;
            const uint32_t nib = ((nz >> 7) * 0x10204080u) >> 28;
            bits |= nib << (4 * k);
        }
    } else {
        for (int k = 0; x0 + k < nx; ++k) bits |= (row[x0 + k] != 0 ? 1u : 0u) << k;
    }
    packed[(int64_t)y * nw + w] = bits;
}

__global__ void life_unpack(const uint32_t *__restrict__ packed, uint8_t *__restrict__ grid, int nx, int ny,
                            int nw, int64_t sx, int64_t origin) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (w >= nw || y >= ny) return;
    uint8_t *row = grid + origin + (int64_t)y * sx;
    const uint32_t bits = packed[(int64_t)y * nw + w];
    const int x0 = 32 * w;
    if (x0 + 32 <= nx) {
        uint32_t q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t nib = (bits >> (4 * k)) & 0xFu;
            q[k] = (nib & 1u) | ((nib & 2u) << 7) | ((nib & 4u) << 14) | ((nib & 8u) << 21);
        }
        *reinterpret_cast<uint4 *>(row + x0) = make_uint4(q[0], q[1], q[2], q[3]);
        *reinterpret_cast<uint4 *>(row + x0 + 16) = make_uint4(q[4], q[5], q[6], q[7]);
    } else {
        for (int k = 0; x0 + k < nx; ++k) row[x0 + k] = (uint8_t)((bits >> k) & 1u);
    }
}

// ---------------------------------------------------------------------------
// bit-packed register temporal blocking
// ---------------------------------------------------------------------------
constexpr int kLanes = 64;
constexpr int kOutCols = kLanes - 2;  // lanes 0 and 63 only feed their neighbours
constexpr int kWavesPerBlock = 4;     // stacked row bands

// word of lane-1 (left neighbour column) and lane+1 (right neighbour column);
// lanes 0 and 63 receive 0 (bound_ctrl) and only ever feed halo columns
template <bool DPP>
__device__ __forceinline__ uint32_t from_left(uint32_t v) {
    if constexpr (DPP) {
        // DPP wave_shr:1 -- lane i reads lane i-1
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xF, 0xF, true);
    } else {
        return (uint32_t)__shfl_up((int)v, 1, kLanes);
    }
}
template <bool DPP>
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
    if constexpr (DPP) {
        // DPP wave_shl:1 -- lane i reads lane i+1
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xF, 0xF, true);
    } else {
        return (uint32_t)__shfl_down((int)v, 1, kLanes);
    }
}

// v_bitop3_b32 truth tables: bit k of the immediate is f(a,b,c) for the
// input bits (a,b,c) of k = 4a + 2b + c, i.e. f evaluated on 0xF0, 0xCC, 0xAA
template <typename F>
constexpr uint32_t truth_table(F f) {
    return f(0xF0u, 0xCCu, 0xAAu) & 0xFFu;
}
constexpr uint32_t kXor3 = truth_table([](uint32_t a, uint32_t b, uint32_t c) { return a ^ b ^ c; });
constexpr uint32_t kMaj3 = truth_table([](uint32_t a, uint32_t b, uint32_t c) { return (a & b) | (a & c) | (b & c); });
// (p1, p0, k1) -> q == 1 and q == 2 where q = p0 + k1 + 2 p1
constexpr uint32_t kQ1 = truth_table([](uint32_t p1, uint32_t p0, uint32_t k1) { return ~p1 & (p0 ^ k1); });
constexpr uint32_t kQ2 = truth_table(
    [](uint32_t p1, uint32_t p0, uint32_t k1) { return (p1 & ~(p0 | k1)) | (~p1 & p0 & k1); });
// (ones, alive, q2) -> even-count survivor term, (ones, q1, t) -> result
constexpr uint32_t kSurvive = truth_table([](uint32_t o, uint32_t a, uint32_t q2) { return ~o & a & q2; });
constexpr uint32_t kResult = truth_table([](uint32_t o, uint32_t q1, uint32_t t) { return (o & q1) | t; });

#define LGD_BITOP3(a, b, c, tt) __builtin_amdgcn_bitop3_b32((a), (b), (c), (tt))

// Horizontal 3-cell sums of one row word: s + 2c = west + centre + east.
template <bool DPP>
__device__ __forceinline__ void row_sums(uint32_t v, uint32_t &s, uint32_t &c) {
    const uint32_t lw = from_left<DPP>(v);
    const uint32_t rw = from_right<DPP>(v);
    const uint32_t west = __builtin_amdgcn_alignbit(v, lw, 31);  // (v << 1) | (lw >> 31)
    const uint32_t east = __builtin_amdgcn_alignbit(rw, v, 1);   // (v >> 1) | (rw << 31)
    s = LGD_BITOP3(west, v, east, kXor3);
    c = LGD_BITOP3(west, v, east, kMaj3);
}

// s9 = (s0 + 2c0) + (s1 + 2c1) + (s2 + 2c2); alive' = s9==3 | (alive & s9==4)
__device__ __forceinline__ uint32_t life_rule(uint32_t s0, uint32_t c0, uint32_t s1, uint32_t c1, uint32_t s2,
                                              uint32_t c2, uint32_t alive) {
    const uint32_t ones = LGD_BITOP3(s0, s1, s2, kXor3);  // weight-1 digit of s9
    const uint32_t k1 = LGD_BITOP3(s0, s1, s2, kMaj3);    // weight-2 carry
    const uint32_t p0 = LGD_BITOP3(c0, c1, c2, kXor3);    // c0+c1+c2 = p0 + 2 p1
    const uint32_t p1 = LGD_BITOP3(c0, c1, c2, kMaj3);
    // q = p0 + k1 + 2 p1 counts the weight-2 units: s9 = ones + 2q
    const uint32_t q1 = LGD_BITOP3(p1, p0, k1, kQ1);      // q == 1
    const uint32_t q2 = LGD_BITOP3(p1, p0, k1, kQ2);      // q == 2
    const uint32_t t = LGD_BITOP3(ones, alive, q2, kSurvive);
    return LGD_BITOP3(ones, q1, t, kResult);
}

// Word column and row a lane loads.  A band reaches `gens` rows and one
// word column past its output; on a wrapped axis (Torus topologies) those are
// the periodic images of grid cells and are computed like them (columns wrap
// word by word, so a wrapped x axis needs nx % 32 == 0), on the others the
// load is clamped into the grid and replaced by the edge state.
__device__ __forceinline__ int packed_col(int col, int nw, int wrap) {
    return (wrap & 1) ? ((col % nw) + nw) % nw : (col < 0 ? 0 : (col >= nw ? nw - 1 : col));
}
__device__ __forceinline__ int packed_row(int row, int ny, int wrap) {
    return (wrap & 2) ? ((row % ny) + ny) % ny : (row < 0 ? 0 : (row >= ny ? ny - 1 : row));
}

template <int R, bool DPP, bool WRAP>
__global__ __launch_bounds__(kLanes *kWavesPerBlock) void life_packed_tb(
    const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, int nw, int ny, uint32_t last_mask,
    uint32_t edge, int gens, int halo_rows, int wrap_mask) {
    // the Cube instantiation (WRAP false) carries no trace of the wrap logic
    const int wrap = WRAP ? wrap_mask : 0;
    const int lane = threadIdx.x & (kLanes - 1);
    // wave index is uniform: keep it (and everything row-related) in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kLanes));
    const int col = (int)blockIdx.x * kOutCols + lane - 1;
    const int band = (int)blockIdx.y * kWavesPerBlock + wave;
    const int valid_rows = R - 2 * halo_rows;
    const int row0 = band * valid_rows - halo_rows;
    if (row0 + halo_rows >= ny) return;  // wave-uniform, no barriers below

    const bool col_in = col >= 0 && col < nw;
    const uint32_t cmask = (wrap & 1) ? 0xFFFFFFFFu : (col_in ? (col == nw - 1 ? last_mask : 0xFFFFFFFFu) : 0u);
    // does this wave touch an unwrapped grid boundary (rows or columns)?
    const bool rows_edge = !(wrap & 2) && (row0 < 0 || row0 + R > ny);
    const bool cols_edge = !(wrap & 1) && (blockIdx.x == 0 || (int)(blockIdx.x + 1) * kOutCols + 1 > nw - 1);
    // register rows [ilo, ihi) lie inside the grid
    int ilo = row0 < 0 ? -row0 : 0;
    int ihi = ny - row0 < R ? ny - row0 : R;

    // every lane loads from an in-bounds column and row (uniform row base +
    // per-lane 32-bit column offset): the periodic image on a wrapped axis,
    // else the clamped cell, replaced by the edge state outside the grid
    const int colc = packed_col(col, nw, wrap);
    uint32_t v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const uint32_t x = src[(int64_t)packed_row(row0 + i, ny, wrap) * nw + colc];
        v[i] = ((wrap & 2) || (i >= ilo && i < ihi)) ? ((x & cmask) | (edge & ~cmask)) : edge;
    }

    for (int g = 0; g < gens; ++g) {
        // keep the row bounds opaque so their per-row tests are evaluated
        // here instead of being hoisted out of the loop as live booleans
        asm volatile("" : "+s"(ilo), "+s"(ihi));
        uint32_t sa, ca, sb, cb;
        row_sums<DPP>(v[0], sa, ca);
        row_sums<DPP>(v[1], sb, cb);
#pragma unroll
        for (int i = 1; i < R - 1; ++i) {
            uint32_t sc, cc;
            row_sums<DPP>(v[i + 1], sc, cc);
            v[i] = life_rule(sa, ca, sb, cb, sc, cc, v[i]);
            sa = sb;
            ca = cb;
            sb = sc;
            cb = cc;
        }
        // cells outside the grid keep the constant edge state
        if (cols_edge) {
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = (v[i] & cmask) | (edge & ~cmask);
        }
        if (rows_edge) {
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = (i >= ilo && i < ihi) ? v[i] : edge;
        }
    }

    if (lane == 0 || lane == kLanes - 1 || !col_in) return;
    const int iend = ihi < R - halo_rows ? ihi : R - halo_rows;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (i >= halo_rows && i < iend) dst[(int64_t)(row0 + i) * nw + col] = v[i];
    }
}

// ---------------------------------------------------------------------------
// stacked waves: NW waves of one work-group hold NW*R consecutive rows of the
// same 64 word columns and hand their boundary rows to each other through
// LDS every generation, so only the work-group's top and bottom K rows are
// redundant halo (instead of 2K rows per wave).
// ---------------------------------------------------------------------------
template <int R, int NW, bool WRAP>
__global__ __launch_bounds__(kLanes *NW) void life_packed_stacked(const uint32_t *__restrict__ src,
                                                                  uint32_t *__restrict__ dst, int nw, int ny,
                                                                  uint32_t last_mask, uint32_t edge, int gens,
                                                                  int halo_rows, int wrap_mask) {
    const int wrap = WRAP ? wrap_mask : 0;
    // [buffer][wave][0: first row, 1: last row][lane]
    __shared__ uint32_t xchg[2][NW][2][kLanes];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kLanes));
    const int col = (int)blockIdx.x * kOutCols + lane - 1;
    const int block_row0 = (int)blockIdx.y * (NW * R - 2 * halo_rows) - halo_rows;
    if (block_row0 + halo_rows >= ny) return;  // block-uniform: no wave is left at a barrier
    const int row0 = block_row0 + wave * R;

    const bool col_in = col >= 0 && col < nw;
    const uint32_t cmask = (wrap & 1) ? 0xFFFFFFFFu : (col_in ? (col == nw - 1 ? last_mask : 0xFFFFFFFFu) : 0u);
    const bool rows_edge = !(wrap & 2) && (row0 < 0 || row0 + R > ny);
    const bool cols_edge = !(wrap & 1) && (blockIdx.x == 0 || (int)(blockIdx.x + 1) * kOutCols + 1 > nw - 1);
    int ilo = row0 < 0 ? (-row0 < R ? -row0 : R) : 0;
    int ihi = ny - row0 < R ? (ny - row0 > 0 ? ny - row0 : 0) : R;

    const int colc = packed_col(col, nw, wrap);
    uint32_t v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const uint32_t x = src[(int64_t)packed_row(row0 + i, ny, wrap) * nw + colc];
        v[i] = ((wrap & 2) || (i >= ilo && i < ihi)) ? ((x & cmask) | (edge & ~cmask)) : edge;
    }

    for (int g = 0; g < gens; ++g) {
        asm volatile("" : "+s"(ilo), "+s"(ihi));
        const int buf = g & 1;
        xchg[buf][wave][0][lane] = v[0];
        xchg[buf][wave][1][lane] = v[R - 1];
        __syncthreads();
        // the block's outermost rows have no neighbour: they decay into the
        // K-row halo like the band edges of life_packed_tb
        const uint32_t above = wave > 0 ? xchg[buf][wave - 1][1][lane] : v[0];
        const uint32_t below = wave < NW - 1 ? xchg[buf][wave + 1][0][lane] : v[R - 1];

        // one generation; the rule is applied to rows [lo, hi) of the wave.
        // The block's top wave has no neighbour above, so after generation g
        // its rows 0 .. g hold don't-care values (halo rows, read only by each
        // other and never stored); likewise rows R-1-g .. R-1 of the bottom
        // wave.  Those rows skip the rule (8 of the 14 operations per row:
        // about 7% of the block's work over a launch; the row sums stay, as
        // skipping them too costs every row a register move).  The bounds
        // are wave-uniform, so a skipped row is a scalar branch.
        auto generation = [&](int lo, int hi) {
            uint32_t sa, ca, sb, cb;
            row_sums<true>(above, sa, ca);
            row_sums<true>(v[0], sb, cb);
#pragma unroll
            for (int i = 0; i < R; ++i) {
                uint32_t sc, cc;
                row_sums<true>(i + 1 < R ? v[i + 1] : below, sc, cc);
                if (i >= lo && i < hi)
                    v[i] = life_rule(sa, ca, sb, cb, sc, cc, v[i]);
                sa = sb;
                ca = cb;
                sb = sc;
                cb = cc;
            }
        };
        generation(wave == 0 ? g + 1 : 0, wave == NW - 1 ? R - 1 - g : R);
        if (cols_edge) {
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = (v[i] & cmask) | (edge & ~cmask);
        }
        if (rows_edge) {
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = (i >= ilo && i < ihi) ? v[i] : edge;
        }
    }

    if (lane == 0 || lane == kLanes - 1 || !col_in) return;
    // rows of this wave inside [block_row0 + K, block_row0 + NW*R - K) and the grid
    const int lo = block_row0 + halo_rows - row0;
    const int hi = block_row0 + NW * R - halo_rows - row0;
    const int iend = ihi < hi ? ihi : hi;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (i >= lo && i < iend) dst[(int64_t)(row0 + i) * nw + col] = v[i];
    }
}

// variant ids of lgdhip_tile::variant for the packed kernel (0 = default)
//   1: independent waves, 128 register rows      2: same with ds_bpermute
//   3: independent waves, 64 register rows       4: same with ds_bpermute
//   5: 4 stacked waves x 64 rows                 6: 8 stacked waves x 64 rows
//   7: 8 stacked waves x 32 rows                 8: 16 stacked waves x 32 rows
constexpr int kNumVariants = 8;
// tuned on MI355X for 16384^2 (8 waves x 32 rows, 31 generations per launch:
// 4.9 ms for 1000 generations); see the variant sweep in the profiles
constexpr int kDefaultVariant = 7;
constexpr int kDefaultPerLaunch = 31;

// register rows one work-group column holds for a variant
int rows_per_group(int variant) {
    switch (variant) {
    case 1: case 2: return 128;
    case 3: case 4: return 64;
    case 5: return 4 * 64;
    case 6: return 8 * 64;
    case 7: return 8 * 32;
    default: return 16 * 32;
    }
}

template <int R>
void launch_rows(const uint32_t *src, uint32_t *dst, int nw, int ny, uint32_t last_mask, uint32_t edge, int gens,
                 int halo_rows, bool dpp, int wrap, hipStream_t s) {
    const int valid = R - 2 * halo_rows;
    const int bands = (ny + valid - 1) / valid;
    dim3 grid((unsigned)((nw + kOutCols - 1) / kOutCols), (unsigned)((bands + kWavesPerBlock - 1) / kWavesPerBlock));
    dim3 block(kLanes * kWavesPerBlock);
    if (dpp && wrap)
        life_packed_tb<R, true, true><<<grid, block, 0, s>>>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, wrap);
    else if (dpp)
        life_packed_tb<R, true, false><<<grid, block, 0, s>>>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, 0);
    else if (wrap)
        life_packed_tb<R, false, true><<<grid, block, 0, s>>>(src, dst, nw, ny, last_mask, edge, gens, halo_rows,
                                                              wrap);
    else
        life_packed_tb<R, false, false><<<grid, block, 0, s>>>(src, dst, nw, ny, last_mask, edge, gens, halo_rows,
                                                               0);
}

template <int R, int NW>
void launch_stacked(const uint32_t *src, uint32_t *dst, int nw, int ny, uint32_t last_mask, uint32_t edge, int gens,
                    int halo_rows, int wrap, hipStream_t s) {
    const int valid = NW * R - 2 * halo_rows;
    dim3 grid((unsigned)((nw + kOutCols - 1) / kOutCols), (unsigned)((ny + valid - 1) / valid));
    if (wrap)
        life_packed_stacked<R, NW, true><<<grid, kLanes * NW, 0, s>>>(src, dst, nw, ny, last_mask, edge, gens,
                                                                      halo_rows, wrap);
    else
        life_packed_stacked<R, NW, false><<<grid, kLanes * NW, 0, s>>>(src, dst, nw, ny, last_mask, edge, gens,
                                                                       halo_rows, 0);
}

int packed_launch(const uint32_t *src, uint32_t *dst, int nw, int ny, uint32_t last_mask, uint32_t edge,
                  int gens, int halo_rows, int variant, int wrap, hipStream_t s) {
    switch (variant) {
    case 1: case 2: launch_rows<128>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, variant == 1, wrap, s); break;
    case 3: case 4: launch_rows<64>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, variant == 3, wrap, s); break;
    case 5: launch_stacked<64, 4>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, wrap, s); break;
    case 6: launch_stacked<64, 8>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, wrap, s); break;
    case 7: launch_stacked<32, 8>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, wrap, s); break;
    default: launch_stacked<32, 16>(src, dst, nw, ny, last_mask, edge, gens, halo_rows, wrap, s); break;
    }
    LGD_LAUNCHED("life_packed_tb");
    return 0;
}

int check_life_grid(const lgdhip_grid *g) {
    int rc = lgdhip::check_grid(g, 2, 1, 1);
    if (rc) return rc;
    if (g->ox % 16 != 0 || g->sx % 16 != 0)
        return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "life kernels need 16-byte aligned rows (ox, sx)");
    // the SWAR and pack kernels read whole 4/32-byte groups past nx
    if (g->sx < g->ox + ((g->nx + 31) / 32) * 32 + 1)
        return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "life kernels need row padding up to a multiple of 32 cells");
    return 0;
}

// K generations per launch over the packed words; wrap bits 1 (x) and 2 (y):
// the axes taken periodic (x only when nx % 32 == 0: columns wrap word by word)
int packed_run(uint32_t *pa, uint32_t *pb, const lgdhip_grid *g, int edge_alive, int steps, int wrap,
               const lgdhip_tile *tile, int *result_in_b, lgdhip_stream_t stream) {
    int rc = lgdhip::check_grid(g, 2, 1, 1);
    if (rc) return rc;
    if ((wrap & 1) && g->nx % 32 != 0)
        return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "a wrapped x axis needs nx % 32 == 0 (whole words)");
    LGD_REQUIRE(pa && pb && pa != pb, "need two distinct device buffers");
    LGD_REQUIRE(steps >= 0, "steps must be >= 0");
    LGD_REQUIRE(edge_alive == 0 || edge_alive == 1, "edge_alive must be 0 or 1");
    int per_launch = kDefaultPerLaunch;
    int variant = kDefaultVariant;
    if (tile) {
        if (tile->steps_per_launch > 0) per_launch = tile->steps_per_launch;
        if (tile->variant != 0) variant = tile->variant;
    }
    if (variant < 1 || variant > kNumVariants) return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "unknown life variant");
    const int rows = rows_per_group(variant);
    // a lane's neighbour word covers 32 cells, so at most 32 generations may
    // run before the halo lanes' garbage reaches an output column; the
    // register rows must keep at least one valid row per band
    if (per_launch < 1 || per_launch > 32 || 2 * per_launch >= rows)
        return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED,
                                 "steps_per_launch must be in [1, 32] and below half the register rows");
    const int nw = (int)((g->nx + 31) / 32);
    const int rem = (int)(g->nx - 32 * (int64_t)(nw - 1));
    const uint32_t last_mask = rem >= 32 ? 0xFFFFFFFFu : ((1u << rem) - 1u);
    const uint32_t edge = edge_alive ? 0xFFFFFFFFu : 0u;
    hipStream_t s = lgdhip::as_stream(stream);
    uint32_t *cur = pa, *nxt = pb;
    for (int done = 0; done < steps;) {
        const int gens = (steps - done < per_launch) ? (steps - done) : per_launch;
        rc = packed_launch(cur, nxt, nw, (int)g->ny, last_mask, edge, gens, per_launch, variant, wrap, s);
        if (rc) return rc;
        uint32_t *tmp = cur;
        cur = nxt;
        nxt = tmp;
        done += gens;
    }
    if (result_in_b) *result_in_b = (cur == pb) ? 1 : 0;
    return 0;
}

} // namespace

extern "C" {

int lgdhip_life2d_u8_run(uint8_t *a, uint8_t *b, const lgdhip_grid *g, int steps, int *result_in_b,
                         lgdhip_stream_t stream) {
    int rc = check_life_grid(g);
    if (rc) return rc;
    LGD_REQUIRE(a && b && a != b, "need two distinct device buffers");
    LGD_REQUIRE(steps >= 0, "steps must be >= 0");
    hipStream_t s = lgdhip::as_stream(stream);
    constexpr int BY = 4;
    dim3 grid(lgdhip::ceil_div((g->nx + 3) / 4, 64), lgdhip::ceil_div(g->ny, BY));
    dim3 block(64, BY);
    const int64_t origin = lgdhip::origin_index(*g);
    uint8_t *cur = a, *nxt = b;
    for (int t = 0; t < steps; ++t) {
        life_u8_swar<BY><<<grid, block, 0, s>>>(cur, nxt, (int)g->nx, (int)g->ny, g->sx, origin);
        LGD_LAUNCHED("life_u8_swar");
        uint8_t *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (result_in_b) *result_in_b = (cur == b) ? 1 : 0;
    return 0;
}

int64_t lgdhip_life2d_packed_words(const lgdhip_grid *g) {
    if (!g) return 0;
    return ((g->nx + 31) / 32) * g->ny;
}

int lgdhip_life2d_pack(const uint8_t *grid, const lgdhip_grid *g, uint32_t *packed, lgdhip_stream_t stream) {
    int rc = check_life_grid(g);
    if (rc) return rc;
    LGD_REQUIRE(grid && packed, "NULL pointer");
    const int nw = (int)((g->nx + 31) / 32);
    LGD_REQUIRE(g->ny <= 65535, "ny must be <= 65535 (one grid row per launch row)");
    dim3 grd(lgdhip::ceil_div(nw, 256), (unsigned)g->ny);
    life_pack<<<grd, 256, 0, lgdhip::as_stream(stream)>>>(grid, packed, (int)g->nx, (int)g->ny, nw, g->sx,
                                                         lgdhip::origin_index(*g));
    LGD_LAUNCHED("life_pack");
    return 0;
}

int lgdhip_life2d_unpack(const uint32_t *packed, const lgdhip_grid *g, uint8_t *grid, lgdhip_stream_t stream) {
    int rc = check_life_grid(g);
    if (rc) return rc;
    LGD_REQUIRE(grid && packed, "NULL pointer");
    LGD_REQUIRE(g->ny <= 65535, "ny must be <= 65535 (one grid row per launch row)");
    const int nw = (int)((g->nx + 31) / 32);
    dim3 grd(lgdhip::ceil_div(nw, 256), (unsigned)g->ny);
    life_unpack<<<grd, 256, 0, lgdhip::as_stream(stream)>>>(packed, grid, (int)g->nx, (int)g->ny, nw, g->sx,
                                                           lgdhip::origin_index(*g));
    LGD_LAUNCHED("life_unpack");
    return 0;
}

int lgdhip_life2d_packed_run(uint32_t *pa, uint32_t *pb, const lgdhip_grid *g, int edge_alive, int steps,
                             const lgdhip_tile *tile, int *result_in_b, lgdhip_stream_t stream) {
    return packed_run(pa, pb, g, edge_alive, steps, 0, tile, result_in_b, stream);
}

int lgdhip_life2d_packed_run_periodic(uint32_t *pa, uint32_t *pb, const lgdhip_grid *g, int edge_alive, int steps,
                                      int wrap_mask, const lgdhip_tile *tile, int *result_in_b,
                                      lgdhip_stream_t stream) {
    LGD_REQUIRE(wrap_mask >= 0 && wrap_mask <= 3, "wrap_mask must be a combination of bits 1 (x) and 2 (y)");
    return packed_run(pa, pb, g, edge_alive, steps, wrap_mask, tile, result_in_b, stream);
}

size_t lgdhip_life2d_workspace_bytes(const lgdhip_grid *g) {
    // two packed buffers, each rounded up to 256 bytes
    const size_t one = (((size_t)lgdhip_life2d_packed_words(g) * 4 + 255) / 256) * 256;
    return 2 * one;
}

int lgdhip_life2d_run(uint8_t *grid, const lgdhip_grid *g, int edge_alive, int steps, void *workspace,
                      size_t workspace_bytes, const lgdhip_tile *tile, lgdhip_stream_t stream) {
    int rc = check_life_grid(g);
    if (rc) return rc;
    LGD_REQUIRE(grid && workspace, "NULL pointer");
    const size_t need = lgdhip_life2d_workspace_bytes(g);
    if (workspace_bytes < need) return lgdhip::set_error(LGDHIP_ERR_WORKSPACE, "life workspace too small");
    uint32_t *pa = (uint32_t *)workspace;
    uint32_t *pb = (uint32_t *)((char *)workspace + need / 2);
    rc = lgdhip_life2d_pack(grid, g, pa, stream);
    if (rc) return rc;
    int in_b = 0;
    rc = lgdhip_life2d_packed_run(pa, pb, g, edge_alive, steps, tile, &in_b, stream);
    if (rc) return rc;
    return lgdhip_life2d_unpack(in_b ? pb : pa, g, grid, stream);
}

} // extern "C"
