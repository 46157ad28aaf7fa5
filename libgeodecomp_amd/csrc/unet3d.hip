// 3D U-Net building blocks on the MFMA cores (BF16 operands, FP32 accumulate).
//
// No reference counterpart (LibGeoDecomp has no learned initial guess).
// Conventions:
//  * activations are channels-last BF16 cubes: element (z,y,x,c) of an n^3
//    tensor with channel stride `ld` and channel offset `off` lives at
//    ((z*n + y)*n + x)*ld + off + c; concatenation is two producers writing
//    disjoint channel ranges of one buffer;
//  * convolution weights are BF16 [cout][Kp] with k = tap*cin + ci,
//    tap = (dz+1)*9 + (dy+1)*3 + (dx+1) for 3x3x3 (tap 0 for 1x1x1), K padded
//    with zeros to Kp = roundup(taps*cin, 32); zero padding at the volume
//    faces ("same" convolution);
//  * a convolution is an implicit GEMM Y[v][co] = sum_k A(v,k) W[co][k] with
//    A(v,k) = X[v + tap(k)][ci(k)], evaluated 16 voxels x 16 channels x 32 k
//    at a time by v_mfma_f32_16x16x32_bf16.
// dgrad is the same kernel with flipped, transposed weights; wgrad reduces
// dY^T X over voxels with the same instruction and a fixed-order reduction.
#include "lgdhip_internal.h"

#include <hip/hip_bf16.h>

#include <algorithm>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

// round-to-nearest-even FP32 -> BF16 (finite values)
__device__ __forceinline__ u16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((uint32_t)h << 16); }

union Frag {
    bf16x8 v;
    uint4 q;
    u16 h[8];
};

__device__ __forceinline__ void tap_offset(int tap, int ksize, int &dz, int &dy, int &dx) {
    if (ksize == 1) {
        dz = dy = dx = 0;
    } else {
        dz = tap / 9 - 1;
        dy = (tap / 3) % 3 - 1;
        dx = tap % 3 - 1;
    }
}

// A fragment: 8 consecutive k of voxel (z, y, x) -- one 16-byte load when
// the 8 k share a tap (cin % 8 == 0); for even cin with even stride and
// offset, 4 loads of two channels (k and k+1 always share a tap there);
// else element-wise.  The fragment's values are the same in every case.
__device__ __forceinline__ Frag load_a(const u16 *__restrict__ X, int n, int ld, int off, int cin, int ksize,
                                       int z, int y, int x, int k0, int K) {
    Frag f;
    f.q = make_uint4(0, 0, 0, 0);
    if ((cin & 7) == 0) {
        if (k0 < K) {
            const int tap = k0 / cin, ci = k0 % cin;
            int dz, dy, dx;
            tap_offset(tap, ksize, dz, dy, dx);
            const int zz = z + dz, yy = y + dy, xx = x + dx;
            if ((unsigned)zz < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)xx < (unsigned)n) {
                f.q = *reinterpret_cast<const uint4 *>(X + ((int64_t)((zz * n + yy) * n + xx)) * ld + off + ci);
            }
        }
    } else if ((cin & 1) == 0 && ((ld | off) & 1) == 0 && ((uintptr_t)X & 3) == 0) {
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            const int k = k0 + e;  // even: channels ci, ci + 1 of one tap
            if (k < K) {
                const int tap = k / cin, ci = k % cin;
                int dz, dy, dx;
                tap_offset(tap, ksize, dz, dy, dx);
                const int zz = z + dz, yy = y + dy, xx = x + dx;
                if ((unsigned)zz < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)xx < (unsigned)n) {
                    const uint32_t pr = *reinterpret_cast<const uint32_t *>(
                        X + ((int64_t)((zz * n + yy) * n + xx)) * ld + off + ci);
                    f.h[e] = (u16)(pr & 0xFFFFu);
                    f.h[e + 1] = (u16)(pr >> 16);
                }
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + e;
            if (k < K) {
                const int tap = k / cin, ci = k % cin;
                int dz, dy, dx;
                tap_offset(tap, ksize, dz, dy, dx);
                const int zz = z + dz, yy = y + dy, xx = x + dx;
                if ((unsigned)zz < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)xx < (unsigned)n)
                    f.h[e] = X[((int64_t)((zz * n + yy) * n + xx)) * ld + off + ci];
            }
        }
    }
    return f;
}

constexpr int kMaxNTiles = 12;  // cout up to 192 (the dgrad of a 192-channel concat input)

// One wave: MT x 16 consecutive voxels (linear order, any n) times NT cout
// tiles of 16, the tiles [g NT, g NT + NT) of cout group g (groups split the
// cout tiles when the voxel tiles alone give too few waves; consecutive waves
// are the groups of one voxel range, so they share its A fragments in
// cache).  Each k step loads MT A fragments and NT B fragments for MT x NT
// MFMAs: the weight fragments are shared by the MT voxel tiles.  Every output
// accumulates over k in the same order whatever the split.
template <int NT, int MT>
__global__ __launch_bounds__(256) void conv3d_fwd_mfma(const u16 *__restrict__ X, int n, int ldx, int offx, int cin,
                                                       int ksize, const u16 *__restrict__ W, int Kp,
                                                       const float *__restrict__ bias, int cout, u16 *Y, int ldy,
                                                       int offy, float *Y32, int ld32, int relu, int accumulate,
                                                       int groups) {
    const int lane = threadIdx.x & 63;
    const int64_t wave_id = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int tile0 = (int)(wave_id % groups) * NT;  // first cout tile of this wave
    const int64_t wave = wave_id / groups;
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v0 = wave * 16 * MT;
    if (v0 >= nvox) return;
    const int i = lane & 15;   // voxel of the A rows / channel of the B columns
    const int kq = lane >> 4;  // which 8 of the 32 k
    const int K = (ksize == 1 ? 1 : 27) * cin;
    // this lane's A voxel in each of the MT tiles
    bool a_ok[MT];
    int ax[MT], ay[MT], az[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int64_t va = v0 + m * 16 + i;
        a_ok[m] = va < nvox;
        ax[m] = (int)(va % n);
        ay[m] = (int)((va / n) % n);
        az[m] = (int)(va / ((int64_t)n * n));
    }

    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int kc = 0; kc < Kp; kc += 32) {
        const int k0 = kc + 8 * kq;
        Frag a[MT], b[NT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            a[m].q = make_uint4(0, 0, 0, 0);
            if (a_ok[m]) a[m] = load_a(X, n, ldx, offx, cin, ksize, az[m], ay[m], ax[m], k0, K);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = (tile0 + t) * 16 + i;
            b[t].q = make_uint4(0, 0, 0, 0);
            if (co < cout) b[t].q = *reinterpret_cast<const uint4 *>(W + (int64_t)co * Kp + k0);
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m].v, b[t].v, acc[m][t], 0, 0, 0);
    }
    // D[r][c] in lane c + 16*(r/4), element r%4: voxel v0 + 16m + 4*(lane/16) + e, channel (tile0+t)*16 + lane%16
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = (tile0 + t) * 16 + i;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t v = v0 + m * 16 + 4 * kq + e;
                if (v >= nvox) continue;
                float r = acc[m][t][e] + bv;
                if (relu) r = r > 0.0f ? r : 0.0f;
                if (Y) {
                    u16 *p = Y + v * ldy + offy + co;
                    if (accumulate) r += bf2f(*p);
                    *p = f2bf(r);
                }
                if (Y32) Y32[v * ld32 + co] = r;
            }
    }
}

// LDS-staged implicit GEMM for 3x3x3 convolutions with at most 32 output
// channels (NT = 1 or 2 tiles of 16): the U-Net's 128^3 and 64^3 levels,
// where a wave's A fragments (16 voxels x 32 k) feed only NT MFMAs each, so
// fetching them from global memory per MFMA is bound by L1 delivery.
// A work-group computes a 16 x 8 x TZ block of output voxels (TZ = 8 / NT);
// each of its 4 waves owns RT = 16 / NT rows of 16 voxels along x.  The
// input channels are taken 8 at a time (a "group"): the block's input tile
// with its one-voxel halo, (16+2) x (8+2) x (TZ+2) voxels x 8 channels, is
// staged in LDS (two buffers: the next group's global loads are in flight
// while the current one is computed), zeros outside the volume giving the
// faces' zero padding.  A group's 27 taps x 8 channels are seven k steps of
// v_mfma_f32_16x16x32_bf16: in step j, lane group q (lanes 16q..16q+15)
// holds tap 4j + q (the 28th tap is zero), so every A fragment is one
// 16-byte LDS read per lane, voxel (x0 + i + dx, y + dy, z + dz).  The
// group's weight fragments (7 x NT) are loaded once and used for all RT
// rows.  The k order (group-major, then tap) differs from
// conv3d_fwd_mfma's; each output is still one FP32 accumulation of the same
// BF16 products.
constexpr int kLdsTX = 16, kLdsTY = 8;
// 16 zero bytes in device memory: the source of loads that would fall
// outside the volume (or past cout), so no load needs a guard or a select.
// Loads go through global-address-space pointers (global_load, counted by
// vmcnt only; a generic pointer would give flat loads, which LDS waits
// would also wait for).
__device__ __align__(16) const unsigned int kZero16[4] = {0u, 0u, 0u, 0u};
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef const u32x4 __attribute__((address_space(1))) *gq_ptr;
typedef const u16 __attribute__((address_space(1))) *gh_ptr;
__device__ __forceinline__ uint4 load_global16(gh_ptr p) {
    const u32x4 v = *(gq_ptr)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
// one 4-byte word (two BF16) where only that word is used: a 16-byte load
// of which the code reads one word would still write all four registers,
// and the compiler, reusing the three it considers dead, then waits for
// the load before going on
typedef const unsigned int __attribute__((address_space(1))) *gw_ptr;
__device__ __forceinline__ unsigned load_global4(gh_ptr p) { return *(gw_ptr)p; }

// IN: how the input is staged -- kInVec, 16-byte loads (cin % 8 == 0);
// kInMasked, 16-byte loads of a row whose channels past cin (< 8) are
// another tensor's or padding and are zeroed; kInElem, one load per
// channel.  VST: the output tile goes out as 16-byte stores through LDS
// (plain BF16 output of full cout tiles, no accumulation, no FP32 copy).
constexpr int kInElem = 0, kInVec = 1, kInMasked = 2;

template <int NT, int IN, bool VST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void conv3d_fwd_lds(const u16 *__restrict__ X, int n, int ldx, int offx, int cin,
                                                      const u16 *__restrict__ W, int Kp,
                                                      const float *__restrict__ bias, int cout, u16 *Y, int ldy,
                                                      int offy, float *Y32, int ld32, int relu, int accumulate,
                                                      int bx, int by) {
    constexpr int TZ = 8 / NT;                // output planes per block
    constexpr int RT = 16 / NT;               // rows per wave: 8 y rows x TZ / 4 planes
    constexpr int SX = kLdsTX + 2, SY = kLdsTY + 2, SZ = TZ + 2;
    constexpr int SV = SX * SY * SZ;          // staged voxels
    constexpr int PER = (SV + 255) / 256;     // staging loads per thread
    __shared__ uint4 tile[2][SV];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, q = lane >> 4;
    int b = blockIdx.x;
    const int x0 = (b % bx) * kLdsTX;
    b /= bx;
    const int y0 = (b % by) * kLdsTY;
    const int z0 = (b / by) * TZ;
    const int groups = (cin + 7) / 8;

    // staged voxel p of this thread: its element offset in X (0 and a zero
    // mask outside the volume, so every load is issued unconditionally and
    // all of a thread's loads are in flight together)
    int64_t soff[PER];
    bool sok[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int sv = tid + 256 * p;
        const int xx = x0 + sv % SX - 1, yy = y0 + (sv / SX) % SY - 1, zz = z0 + sv / (SX * SY) - 1;
        sok[p] = sv < SV && (unsigned)xx < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)zz < (unsigned)n;
        soff[p] = sok[p] ? (((int64_t)zz * n + yy) * n + xx) * ldx : 0;
    }
    uint4 stage[PER];
    const gh_ptr zero = (gh_ptr)(const void *)kZero16;
    const gh_ptr xg = (gh_ptr)X;
    auto stage_load = [&](int c) {
        const gh_ptr base = xg + offx + 8 * c;
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            Frag f;
            if (IN != kInElem) {
                f.q = load_global16(sok[p] ? base + soff[p] : zero);
                if (IN == kInMasked) {
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        if (e >= cin) f.h[e] = 0;
                }
            } else {
                // channels beyond cin (and voxels outside) read zeros
#pragma unroll
                for (int e = 0; e < 8; ++e) f.h[e] = *(sok[p] && 8 * c + e < cin ? base + soff[p] + e : zero);
            }
            stage[p] = f.q;
        }
    };
    auto stage_store = [&](int buf) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int sv = tid + 256 * p;
            if (sv < SV) tile[buf][sv] = stage[p];
        }
    };

    // this lane group's tap of each k step, as an offset in the staged tile
    // relative to the row's first voxel (x0 - 1, y - 1, z - 1)
    int toff[7];
    bool tok[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const int tap = 4 * j + q;
        tok[j] = tap < 27;
        int dz, dy, dx;
        tap_offset(tok[j] ? tap : 13, 3, dz, dy, dx);
        toff[j] = ((dz + 1) * SY + (dy + 1)) * SX + (dx + 1) + i;
    }

    f32x4 acc[RT][NT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    // a group's weight fragments: channels 8c..8c+7 of tap 4j + q for output
    // channel t * 16 + i (zero for the 28th tap, channels past cout or cin)
    Frag bf[7][NT];
    auto load_b = [&](int c) {
#pragma unroll
        for (int j = 0; j < 7; ++j)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int co = t * 16 + i;
                const bool ok = tok[j] && co < cout;
                const gh_ptr src = (gh_ptr)W + (int64_t)co * Kp + (4 * j + q) * cin + 8 * c;
                if (IN == kInVec) {
                    bf[j][t].q = load_global16(ok ? src : zero);
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) bf[j][t].h[e] = *(ok && 8 * c + e < cin ? src + e : zero);
                }
            }
    };

    // Software pipeline over the groups.  Loads complete in issue order
    // (vmcnt), so each group's loads are issued after the ones it waits for
    // last: the staged input of group c + 1 goes out first, the MFMAs of
    // group c then need only the weights loaded in the previous iteration;
    // the weights of group c + 1 go out after the MFMAs, and writing the
    // staged input waits only for the older staging loads.
    stage_load(0);
    load_b(0);
    stage_store(0);
    __syncthreads();
    for (int c = 0; c < groups; ++c) {
        const int cur = c & 1;
        // the next group's loads are issued unconditionally (the last group
        // re-reads its own, never stored), so the wait counts are the same on
        // every iteration and exact
        const int cn = c + 1 < groups ? c + 1 : c;
        stage_load(cn);  // in flight during the MFMAs below
        // k step outermost: consecutive MFMAs update different rows'
        // accumulators, so none waits for the previous one's result
        // (the 28th tap's zero weights make its A values irrelevant: it reads
        // the block's centre tap, always inside the staged tile)
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            Frag a[RT];  // all rows' fragments of this k step in flight together
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                const int yl = r % 8, zl = wave * (TZ / 4) + r / 8;
                a[r].q = tile[cur][(zl * SY + yl) * SX + toff[j]];
            }
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r].v, bf[j][t].v, acc[r][t], 0, 0, 0);
        }
        // schedule: kLdsAhead fragment reads in flight ahead of the MFMAs
        // that consume them (the compiler alone keeps two, and each MFMA
        // then waits out an LDS latency)
        constexpr int kReads = 7 * RT, kLdsAhead = 8;
        __builtin_amdgcn_sched_group_barrier(0x100, kLdsAhead, 0);
        lgdhip::static_for<0, kReads - kLdsAhead>([&](auto) {
            __builtin_amdgcn_sched_group_barrier(0x008, NT, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        });
        __builtin_amdgcn_sched_group_barrier(0x008, NT * kLdsAhead, 0);
        load_b(cn);
        if (c + 1 < groups) stage_store(cur ^ 1);
        __syncthreads();
    }
    // D[r][c] in lane c + 16 (r / 4), element r % 4: voxel x0 + 4q + e of the
    // row, channel t * 16 + i
    if (VST) {
        // through LDS (the staging buffers are free after the last barrier):
        // [voxel of the block][16 NT channels] BF16, then 16-byte stores
        constexpr int CH = 16 * NT;
        u16 *otile = reinterpret_cast<u16 *>(&tile[0][0]);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + i;
            const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                const int vt = ((wave * (TZ / 4) + r / 8) * 8 + r % 8) * 16 + 4 * q;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float o = acc[r][t][e] + bv;
                    if (relu) o = o > 0.0f ? o : 0.0f;
                    otile[(vt + e) * CH + co] = f2bf(o);
                }
            }
        }
        __syncthreads();
        constexpr int CHUNKS = 16 * 8 * TZ * CH / 8;  // 16-byte pieces of the block
#pragma unroll
        for (int k = tid; k < CHUNKS; k += 256) {
            const int vt = k / (CH / 8), part = k % (CH / 8);
            const int x = x0 + vt % 16, y = y0 + (vt / 16) % 8, z = z0 + vt / 128;
            if (x < n && y < n && z < n) {
                const int64_t v = ((int64_t)z * n + y) * n + x;
                *reinterpret_cast<uint4 *>(Y + v * ldy + offy + 8 * part) =
                    *reinterpret_cast<const uint4 *>(otile + vt * CH + 8 * part);
            }
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + i;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int y = y0 + r % 8, z = z0 + wave * (TZ / 4) + r / 8;
            if (y >= n || z >= n) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int x = x0 + 4 * q + e;
                if (x >= n) continue;
                const int64_t v = ((int64_t)z * n + y) * n + x;
                float o = acc[r][t][e] + bv;
                if (relu) o = o > 0.0f ? o : 0.0f;
                if (Y) {
                    u16 *pp = Y + v * ldy + offy + co;
                    if (accumulate) o += bf2f(*pp);
                    *pp = f2bf(o);
                }
                if (Y32) Y32[v * ld32 + co] = o;
            }
        }
    }
}

// Z-marching LDS convolution for 3x3x3 layers with up to 16 output channels
// and cin = 8 G input channels (G <= 6; G = 1 also takes cin < 8 rows, see
// IN).  A work-group owns a 16 x 8 column of output voxels and marches z
// over a chunk of ZC planes.  Input planes (18 x 10 voxels with their halo,
// all cin channels) pass through a ring of 4 LDS slots: output plane z reads
// input planes z-1, z, z+1 while plane z+2, loaded into registers one step
// earlier, is written into the fourth slot -- every input voxel is fetched
// once per column, whole channel rows together (the tile kernel above takes
// the channels 8 at a time, re-reading each voxel's row G times).  The G x 7
// weight fragments of the lane stay in registers for the whole march; wave w
// computes output rows y0 + 2w and y0 + 2w + 1 of each plane.  The output
// plane goes out through LDS as 16-byte stores one step later.  k order as
// conv3d_fwd_lds (group-major, then tap).
constexpr int kZmChunk = 32;  // output planes per work-group (16: forward 731 us, 32: 705 us at 128^3; 8: 782 us)

template <int G, int IN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void conv3d_fwd_zm(
    const u16 *__restrict__ X, int n, int ldx, int offx, int cin, const u16 *__restrict__ W, int Kp,
    const float *__restrict__ bias, int cout, u16 *Y, int ldy, int offy, int relu, int bx, int by) {
    constexpr int SX = 18, SY = 10, PV = SX * SY;  // staged voxels per plane
    constexpr int CHUNKS = PV * G;                 // 16-byte pieces per plane
    constexpr int PER = (CHUNKS + 255) / 256;
    __shared__ uint4 ring[4][CHUNKS];               // [slot][group][voxel]
    __shared__ u16 oplane[2][8 * 16 * 16];          // [y][x][channel] of an output plane
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, q = lane >> 4;
    int b = blockIdx.x;
    const int x0 = (b % bx) * 16;
    b /= bx;
    const int y0 = (b % by) * 8;
    const int z0 = (b / by) * kZmChunk;
    const int zend = min(z0 + kZmChunk, n);

    const gh_ptr zero = (gh_ptr)(const void *)kZero16;
    const gh_ptr xg = (gh_ptr)X + offx;
    // this thread's pieces of a plane: voxel sv = k / G, group g = k % G
    // (consecutive threads read consecutive 16 bytes of the voxels' rows)
    int pofs[PER];   // element offset within a plane's (x, y) footprint
    bool pok[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int k = tid + 256 * p;
        const int sv = k / G, g = k % G;
        const int xx = x0 + sv % SX - 1, yy = y0 + sv / SX - 1;
        pok[p] = k < CHUNKS && (unsigned)xx < (unsigned)n && (unsigned)yy < (unsigned)n;
        pofs[p] = pok[p] ? (yy * n + xx) * ldx + 8 * g : 0;
    }
    uint4 stage[PER];
    auto plane_load = [&](int z) {
        const bool zok = (unsigned)z < (unsigned)n;
        const gh_ptr base = xg + (int64_t)(zok ? z : 0) * n * n * ldx;
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            Frag f;
            f.q = load_global16(pok[p] && zok ? base + pofs[p] : zero);
            if (IN == kInMasked) {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (e >= cin) f.h[e] = 0;
            }
            stage[p] = f.q;
        }
    };
    auto plane_store = [&](int slot) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int k = tid + 256 * p;
            if (k < CHUNKS) ring[slot][(k % G) * PV + k / G] = stage[p];
        }
    };

    // weights: group g, k step j -> tap 4j + q, channels 8g..8g+7 of output
    // channel i (zero for the 28th tap and past cout / cin)
    Frag bf[G][7];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const bool ok = 4 * j + q < 27 && i < cout;
            const gh_ptr src = (gh_ptr)W + (int64_t)i * Kp + (4 * j + q) * cin + 8 * g;
            if (IN == kInVec) {
                bf[g][j].q = load_global16(ok ? src : zero);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) bf[g][j].h[e] = *(ok && e < cin ? src + e : zero);
            }
        }
    // this lane group's taps: ring slot offset (dz + 1), row / column offset
    int tslot[7], toff[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const int tap = 4 * j + q;
        int dz, dy, dx;
        tap_offset(tap < 27 ? tap : 13, 3, dz, dy, dx);
        tslot[j] = dz + 1;
        toff[j] = (dy + 1) * SX + (dx + 1) + i;
    }
    const float bv = bias && i < cout ? bias[i] : 0.0f;

    // prologue: planes z0-1, z0, z0+1 in slots 0..2, z0+2 in flight
    plane_load(z0 - 1);
    plane_store(0);
    plane_load(z0);
    plane_store(1);
    plane_load(z0 + 1);
    plane_store(2);
    plane_load(z0 + 2);
    __syncthreads();
    for (int z = z0, st = 0; z < zend; ++z, ++st) {
        // the previous plane's output tile goes out (written before the barrier)
        if (st > 0) {
            const int ob = (st - 1) & 1;
            const int vt = tid >> 1, part = tid & 1;  // 128 voxels x 2 halves of 16 channels
            const int x = x0 + vt % 16, y = y0 + vt / 16;
            if (x < n && y < n)
                *reinterpret_cast<uint4 *>(Y + (((int64_t)(z - 1) * n + y) * n + x) * ldy + offy + 8 * part) =
                    *reinterpret_cast<const uint4 *>(&oplane[ob][vt * 16 + 8 * part]);
        }
        f32x4 acc[2];
        acc[0] = acc[1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const uint4 *pl = &ring[(st + tslot[j]) & 3][g * PV];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    Frag a;
                    a.q = pl[(2 * wave + r) * SX + toff[j]];
                    acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, bf[g][j].v, acc[r], 0, 0, 0);
                }
            }
        // output plane z into LDS: row 2w + r, voxels 4q + e, channel i
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float o = acc[r][e] + bv;
                if (relu) o = o > 0.0f ? o : 0.0f;
                oplane[st & 1][((2 * wave + r) * 16 + 4 * q + e) * 16 + i] = f2bf(o);
            }
        // plane z+2 into the free slot, plane z+3 in flight
        plane_store((st + 3) & 3);
        plane_load(z + 3);
        __syncthreads();
    }
    // the last plane's output
    {
        const int st = zend - z0;
        const int ob = (st - 1) & 1;
        const int vt = tid >> 1, part = tid & 1;
        const int x = x0 + vt % 16, y = y0 + vt / 16;
        if (x < n && y < n)
            *reinterpret_cast<uint4 *>(Y + (((int64_t)(zend - 1) * n + y) * n + x) * ldy + offy + 8 * part) =
                *reinterpret_cast<const uint4 *>(&oplane[ob][vt * 16 + 8 * part]);
    }
}

// GEMM-tiled implicit convolution for 3x3x3 layers with cin % 8 == 0 on the
// small levels (n < 64) and for cout > 32: Y[v][co] = sum_k A(v, k) W[co][k]
// over a 64-voxel x 64-channel block per work-group.  k runs in stages of 64
// (eight 8-channel chunks of one or more taps); a stage's A block (64 voxels
// x 64 k, gathered from the channels-last input with zeros outside the
// volume) and B block (64 channels x 64 k of the weights) are staged in LDS,
// double buffered: the next stage's global loads are in flight in registers
// while the current one is computed.  Wave w computes the 32 x 32 quarter
// (voxel half w & 1, channel half w >> 1) as 2 x 2 MFMA tiles, two k steps
// of 32 per stage, every fragment one ds_read_b128 (rows padded to 72
// elements: the 16 rows of a fragment read hit distinct banks).  A block
// reads its A and B data from L2 once instead of once per wave.
// With `slices` > 1 the stage range is split over work-groups (a slice
// each, for small volumes whose 64 x 64 blocks alone leave CUs idle) and
// every slice writes FP32 partials P[slice][v][co]; conv3d_gemm_reduce adds
// them in slice order.  k order: stage-major, then k within the stage (the
// tap-major order of the weights), per slice.
constexpr int kGemmBM = 64, kGemmBN = 64, kGemmBK = 64;
constexpr int kGemmLd = kGemmBK + 8;  // LDS row stride in elements (144 bytes)

template <bool SPLIT>
__global__ __launch_bounds__(256) void conv3d_fwd_gemm(const u16 *__restrict__ X, int n, int ldx, int offx, int cin,
                                                       const u16 *__restrict__ W, int Kp,
                                                       const float *__restrict__ bias, int cout, u16 *Y, int ldy,
                                                       int offy, float *Y32, int ld32, int relu, int accumulate,
                                                       int mblocks, int nblocks, int slices, int stages,
                                                       float *__restrict__ P) {
    __shared__ uint4 lds[2][2][kGemmBM * kGemmLd / 8];  // [buffer][A, B][row * kGemmLd / 8 + chunk]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, q = lane >> 4;
    int b = blockIdx.x;
    const int mb = b % mblocks;
    b /= mblocks;
    const int nb = b % nblocks;
    const int slice = b / nblocks;
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v0 = (int64_t)mb * kGemmBM;
    const int co0 = nb * kGemmBN;
    const int st0 = (int)((int64_t)slice * stages / slices), st1 = (int)((int64_t)(slice + 1) * stages / slices);
    const int K = 27 * cin;

    // loader roles: chunk c = tid + 256 u (u = 0, 1) of the 512 16-byte
    // chunks of each block: row c / 8, k chunk c % 8 (the same for both)
    const int j = tid & 7;
    const gh_ptr zero = (gh_ptr)(const void *)kZero16;
    const gh_ptr xg = (gh_ptr)X + offx;
    int vx[2], vy[2], vz[2];
    bool vok[2];
    const gh_ptr wrow[2] = {(gh_ptr)W + (int64_t)(co0 + (tid >> 3)) * Kp,
                            (gh_ptr)W + (int64_t)(co0 + (tid >> 3) + 32) * Kp};
    const bool cok[2] = {co0 + (tid >> 3) < cout, co0 + (tid >> 3) + 32 < cout};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int64_t v = v0 + (tid >> 3) + 32 * u;
        vok[u] = v < nvox;
        vx[u] = (int)(v % n);
        vy[u] = (int)((v / n) % n);
        vz[u] = (int)(v / ((int64_t)n * n));
    }
    const float inv_cin = 1.0f / (float)cin;
    uint4 ra[2], rb[2];
    auto load_stage = [&](int st) {
        const int k0 = st * kGemmBK + 8 * j;
        // tap = k0 / cin from a float reciprocal, corrected (k0 < 2^24)
        int tap = (int)((float)k0 * inv_cin);
        tap -= tap * cin > k0;
        tap += (tap + 1) * cin <= k0;
        const int ci = k0 - tap * cin;
        const bool kin = k0 < K;
        int dz, dy, dx;
        tap_offset(kin ? tap : 13, 3, dz, dy, dx);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int zz = vz[u] + dz, yy = vy[u] + dy, xx = vx[u] + dx;
            const bool ok = kin && vok[u] && (unsigned)zz < (unsigned)n && (unsigned)yy < (unsigned)n &&
                            (unsigned)xx < (unsigned)n;
            ra[u] = load_global16(ok ? xg + ((int64_t)(zz * n + yy) * n + xx) * ldx + ci : zero);
            rb[u] = load_global16(cok[u] && k0 < Kp ? wrow[u] + k0 : zero);
        }
    };
    auto store_stage = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int row = (tid >> 3) + 32 * u;
            lds[buf][0][row * (kGemmLd / 8) + j] = ra[u];
            lds[buf][1][row * (kGemmLd / 8) + j] = rb[u];
        }
    };

    f32x4 acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int wm = wave & 1, wn = wave >> 1;
    if (st0 < st1) {
        load_stage(st0);
        store_stage(0);
        __syncthreads();
        for (int st = st0; st < st1; ++st) {
            const int cur = (st - st0) & 1;
            // the next stage's loads go out first (the last stage re-loads
            // its own, never stored, so every iteration waits alike)
            load_stage(st + 1 < st1 ? st + 1 : st);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                Frag a[2], bf[2];
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    a[m].q = lds[cur][0][(32 * wm + 16 * m + i) * (kGemmLd / 8) + 4 * ks + q];
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    bf[t].q = lds[cur][1][(32 * wn + 16 * t + i) * (kGemmLd / 8) + 4 * ks + q];
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m].v, bf[t].v, acc[m][t], 0, 0, 0);
            }
            if (st + 1 < st1) store_stage(cur ^ 1);
            __syncthreads();
        }
    }
    // D[r][c] in lane c + 16 (r / 4), element r % 4: voxel row 4q + e of the
    // tile, channel i
    const int ldp = nblocks * kGemmBN;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int co = co0 + 32 * wn + 16 * t + i;
        const float bv = !SPLIT && bias && co < cout ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t v = v0 + 32 * wm + 16 * m + 4 * q + e;
                if (v >= nvox) continue;
                if (SPLIT) {
                    P[((int64_t)slice * nvox + v) * ldp + co] = acc[m][t][e];
                    continue;
                }
                if (co >= cout) continue;
                float r = acc[m][t][e] + bv;
                if (relu) r = r > 0.0f ? r : 0.0f;
                if (Y) {
                    u16 *pp = Y + v * ldy + offy + co;
                    if (accumulate) r += bf2f(*pp);
                    *pp = f2bf(r);
                }
                if (Y32) Y32[v * ld32 + co] = r;
            }
    }
}

// the slices' partials of conv3d_fwd_gemm summed in slice order, then bias,
// ReLU, accumulation and the BF16 / FP32 stores of the unsplit epilogue
__global__ __launch_bounds__(256) void conv3d_gemm_reduce(const float *__restrict__ P, int64_t nvox, int ldp,
                                                          int slices, const float *__restrict__ bias, int cout,
                                                          u16 *Y, int ldy, int offy, float *Y32, int ld32, int relu,
                                                          int accumulate) {
    // a split launch has fewer than 1024 blocks of 64 x 64 outputs, so
    // nvox * cout < 2^22 and 32-bit index arithmetic holds
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= (uint32_t)(nvox * cout)) return;
    const uint32_t v = idx / (uint32_t)cout;
    const int co = (int)(idx - v * (uint32_t)cout);
    float r = 0.0f;
    for (int sl = 0; sl < slices; ++sl) r += P[((int64_t)sl * nvox + v) * ldp + co];
    r += bias ? bias[co] : 0.0f;
    if (relu) r = r > 0.0f ? r : 0.0f;
    if (Y) {
        u16 *pp = Y + (int64_t)v * ldy + offy + co;
        if (accumulate) r += bf2f(*pp);
        *pp = f2bf(r);
    }
    if (Y32) Y32[(int64_t)v * ld32 + co] = r;
}

// 1x1x1 convolution to a single output channel (the U-Net's output layer):
// one thread per voxel, its cin channels in 16-byte loads (cin % 8 == 0), an
// FP32 dot product with the cin weights.  An MFMA tile would compute 16
// output channels to keep one, and the layer is memory bound.
__global__ __launch_bounds__(256) void conv1x1_to1(const u16 *__restrict__ X, int64_t nvox, int ldx, int offx,
                                                   int cin, const u16 *__restrict__ W,
                                                   const float *__restrict__ bias, u16 *Y, int ldy, int offy,
                                                   float *Y32, int ld32, int relu, int accumulate) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= nvox) return;
    const u16 *xr = X + v * ldx + offx;
    float acc = 0.0f;
    for (int c = 0; c < cin; c += 8) {
        Frag f, w;
        f.q = *reinterpret_cast<const uint4 *>(xr + c);
        w.q = *reinterpret_cast<const uint4 *>(W + c);
#pragma unroll
        for (int e = 0; e < 8; ++e) acc += bf2f(f.h[e]) * bf2f(w.h[e]);
    }
    float o = acc + (bias ? bias[0] : 0.0f);
    if (relu) o = o > 0.0f ? o : 0.0f;
    if (Y) {
        u16 *pp = Y + v * ldy + offy;
        if (accumulate) o += bf2f(*pp);
        *pp = f2bf(o);
    }
    if (Y32) Y32[v * ld32] = o;
}

// the LDS-staged kernel takes 3x3x3 layers with up to 32 output channels on
// volumes of at least 16^3
constexpr int kLdsMinN = 16;
// z-marching in 16-channel slices for wider outputs from this edge on (96^3:
// 432 work-groups per slice launch; 64^3 would leave half the CUs idle)
constexpr int kZmSliceMinN = 96;

// The GEMM-tiled kernel's launch: 64 x 64 blocks, stages of 64 k, and the
// number of slices the stage range is split into when a workspace for the
// partials is given -- enough work-groups for 4 per CU, at least 4 stages
// per slice.
struct GemmPlan {
    int mblocks, nblocks, stages, slices;
};
GemmPlan gemm_plan(int n, int cin, int cout, bool split) {
    GemmPlan p;
    const int64_t nvox = (int64_t)n * n * n;
    p.mblocks = (int)((nvox + kGemmBM - 1) / kGemmBM);
    p.nblocks = (cout + kGemmBN - 1) / kGemmBN;
    p.stages = (lgdhip_conv3d_kpad(cin, 3) + kGemmBK - 1) / kGemmBK;
    p.slices = 1;
    const int64_t base = (int64_t)p.mblocks * p.nblocks, target = 4 * lgdhip::kCUs;
    if (split && base < target) {
        const int64_t want = (target + base - 1) / base;
        p.slices = (int)std::max<int64_t>(1, std::min<int64_t>(want, p.stages / 4));
    }
    return p;
}
size_t gemm_workspace_bytes(int n, const GemmPlan &p) {
    if (p.slices <= 1) return 0;
    return (size_t)p.slices * (size_t)n * n * n * (size_t)p.nblocks * kGemmBN * 4;
}
// the layers the GEMM-tiled kernel takes (16-byte input rows; the caller
// has checked ldx, offx and the alignment when cin % 8 == 0): 3x3x3 with
// cin % 8 == 0 and more than 16 output channels below 64^3, and any size
// with more than 32 (the LDS-tile kernel takes 32 channels from 64^3 on;
// 16-channel layers stay with the z-marching kernel and, for the launches it
// does not take, the LDS-tile kernel, which sums in the same k order)
bool gemm_applies(int n, int cin, int ksize, int cout) {
    return ksize == 3 && (cin & 7) == 0 && cout > 16 && (n < 64 || cout > 32);
}

// voxel tiles per wave of the forward kernel: weight-fragment reuse bounded
// by the accumulator registers (MT x NT x 4)
constexpr int fwd_mt(int nt) { return nt <= 4 ? 4 : (nt <= 8 ? 2 : 1); }

// Forward launch shape for nvox voxels and nt cout tiles: NT tiles per wave
// (a divisor of nt), MT voxel tiles per wave (at most fwd_mt(NT)), nt / NT
// cout groups.  Takes the fewest fragment loads per MFMA (1/NT + 1/MT) among
// the shapes that give at least kFwdWaves waves (8 per CU: the 16^3 and 32^3
// levels of the U-Net otherwise leave most CUs idle), else the most waves.
constexpr int64_t kFwdWaves = 2048;
struct FwdShape {
    int nt, mt, groups;
};
FwdShape fwd_shape(int64_t nvox, int nt) {
    FwdShape best{nt, fwd_mt(nt), 1}, most = best;
    double best_loads = 1e30;
    int64_t most_waves = -1;
    for (int ntw = nt; ntw >= 1; --ntw) {
        if (nt % ntw) continue;
        for (int mt = fwd_mt(ntw); mt >= 1; mt /= 2) {
            const int64_t waves = (nvox + 16 * mt - 1) / (16 * mt) * (nt / ntw);
            const double loads = 1.0 / ntw + 1.0 / mt;
            if (waves >= kFwdWaves && loads < best_loads) {
                best_loads = loads;
                best = FwdShape{ntw, mt, nt / ntw};
            }
            if (waves > most_waves) {
                most_waves = waves;
                most = FwdShape{ntw, mt, nt / ntw};
            }
        }
    }
    return best_loads < 1e30 ? best : most;
}

// dW partial over a chunk of voxels: P[chunk][co][k] = sum_v dY[v][co] X[v+tap(k)][ci(k)]
// one wave per (chunk, k tile of 16), all co tiles; work-groups in XCD
// order (XCD c takes the contiguous range [c per, (c + 1) per): a chunk's
// waves and its neighbours share one L2)
template <int NT>
__global__ __launch_bounds__(256) void conv3d_wgrad_mfma(const u16 *__restrict__ X, int n, int ldx, int offx, int cin,
                                                         int ksize, const u16 *__restrict__ dY, int ldy, int offy,
                                                         int cout, int Kp, int chunk_vox, float *__restrict__ P,
                                                         unsigned per) {
    const int lane = threadIdx.x & 63;
    const int ktiles = Kp / 16;
    const int wave = (int)(lgdhip::xcd_tile(blockIdx.x, per, true) * 4 + threadIdx.x / 64);
    const int nvox = n * n * n;
    const int chunks = (nvox + chunk_vox - 1) / chunk_vox;
    if (wave >= chunks * ktiles) return;
    const int kt = wave % ktiles;
    const int chunk = wave / ktiles;
    const int i = lane & 15;
    const int vq = lane >> 4;
    const int K = (ksize == 1 ? 1 : 27) * cin;
    const int k = kt * 16 + i;  // this lane's B column
    int tap = 0, ci = 0, dz = 0, dy = 0, dx = 0;
    if (k < K) {
        tap = k / cin;
        ci = k % cin;
        tap_offset(tap, ksize, dz, dy, dx);
    }
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int v_begin = chunk * chunk_vox;
    const int v_end = min(v_begin + chunk_vox, nvox);
    // coordinates of this lane group's first voxel vb + 8 vq, carried along
    // (one voxel per element, 32 per iteration) instead of divided out per
    // element
    int x0 = (v_begin + 8 * vq) % n, y0 = ((v_begin + 8 * vq) / n) % n, z0 = (v_begin + 8 * vq) / (n * n);
    auto advance = [n](int &xa, int &ya, int &za, int by) {
        xa += by;
        while (xa >= n) {
            xa -= n;
            if (++ya == n) {
                ya = 0;
                ++za;
            }
        }
    };
    for (int vb = v_begin; vb < v_end; vb += 32) {
        Frag b;
        int x = x0, y = y0, z = z0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int v = vb + 8 * vq + e;
            u16 val = 0;
            if (v < v_end && k < K) {
                const int zz = z + dz, yy = y + dy, xx = x + dx;
                if ((unsigned)zz < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)xx < (unsigned)n)
                    val = X[((int64_t)((zz * n + yy) * n + xx)) * ldx + offx + ci];
            }
            b.h[e] = val;
            if (++x == n) {  // the next voxel
                x = 0;
                if (++y == n) {
                    y = 0;
                    ++z;
                }
            }
        }
        advance(x0, y0, z0, 32);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            Frag a;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int v = vb + 8 * vq + e;
                const int co = t * 16 + i;
                a.h[e] = (v < v_end && co < cout) ? dY[(int64_t)v * ldy + offy + co] : (u16)0;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[t], 0, 0, 0);
        }
    }
    // D[co][k]: lane c + 16*(r/4) holds row 4*(lane/16)+e (co), column lane%16 (k)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int co = t * 16 + 4 * vq + e;
            if (co < cout) P[((int64_t)chunk * cout + co) * Kp + kt * 16 + i] = acc[t][e];
        }
    }
}

// channels-last (ld, off, C) -> voxel-blocked [nvox / 8][C][8] (C, ld, off
// and nvox multiples of 8): element (v, c) at ((v / 8) C + c) 8 + v % 8, so
// a channel's 8 consecutive voxels are 16 contiguous bytes and the 16
// channels of an MFMA fragment at one 8-voxel group 256 contiguous bytes.
// A work-group takes 256 voxels and 8 channels: coalesced 16-byte reads of
// each voxel's 8 channels, a transpose through LDS, 16-byte writes of each
// channel's 8 voxels (128 contiguous bytes per 8 lanes).
__global__ __launch_bounds__(256) void to_voxel_blocked(const u16 *__restrict__ src, int64_t nvox, int ld, int off,
                                                        int C, u16 *__restrict__ dst) {
    __shared__ uint4 tile[256];  // [voxel of the block][8 channels]
    const int64_t v0 = (int64_t)blockIdx.x * 256;
    const int g = blockIdx.y;  // channels 8g .. 8g + 7
    const int t = threadIdx.x;
    if (v0 + t < nvox) tile[t] = *(const uint4 *)(src + (v0 + t) * ld + off + 8 * g);
    __syncthreads();
    const int vg = t >> 3, c = t & 7;  // voxel group of the block, channel
    const int64_t vb = v0 + 8 * vg;
    if (vb >= nvox) return;  // after the only barrier; nvox % 8 == 0
    const u16 *tl = reinterpret_cast<const u16 *>(tile);
    Frag f;
#pragma unroll
    for (int e = 0; e < 8; ++e) f.h[e] = tl[(8 * vg + e) * 8 + c];
    *reinterpret_cast<uint4 *>(dst + (((vb >> 3) * C + 8 * g + c) << 3)) = f.q;
}

// 8 BF16 of one channel's voxels shifted by one voxel: dx = -1 gives
// element e = c[e-1] with element 0 from the previous 8 (their last word,
// elements 6 and 7: prev_w); dx = +1 gives c[e+1] with element 7 from the
// next 8 (their first word, elements 0 and 1: next_x); little-endian
// halves, so each 32-bit word is an alignbit of two neighbours
__device__ __forceinline__ uint4 shift_from_prev(uint4 c, unsigned prev_w) {
    return make_uint4(__builtin_amdgcn_alignbit(c.x, prev_w, 16), __builtin_amdgcn_alignbit(c.y, c.x, 16),
                      __builtin_amdgcn_alignbit(c.z, c.y, 16), __builtin_amdgcn_alignbit(c.w, c.z, 16));
}
__device__ __forceinline__ uint4 shift_from_next(uint4 c, unsigned next_x) {
    return make_uint4(__builtin_amdgcn_alignbit(c.y, c.x, 16), __builtin_amdgcn_alignbit(c.z, c.y, 16),
                      __builtin_amdgcn_alignbit(c.w, c.z, 16), __builtin_amdgcn_alignbit(next_x, c.w, 16));
}

// dW partial from voxel-blocked copies: P[chunk][co][tap*cin + ci] =
// sum over the chunk's voxels v of G[co][v] X[ci][v + tap].  One wave per
// (chunk, taps, group of CT co tiles, group of IT ci tiles).  A wave's
// taps: DXT = 3 takes the three x-shifted taps of a (dz, dy) row, whose B
// fragments come from the row's centre 8 voxels and one word of each
// neighbouring 8; DYT = 3 (with DXT = 3) takes the three rows of one dz as
// well, so 9 taps share the A fragment, the voxel stepping and the
// validity tests (per MFMA about a third of the bookkeeping of one row);
// DXT = 1: one tap.  Per 32 voxels (4 groups of 8 consecutive voxels of one
// x row, n % 8 == 0) every fragment is one 16-byte load per lane, the 16
// lanes of a voxel group reading 256 contiguous bytes of the voxel-blocked
// copy.  Needs ksize 3, cin and cout multiples of 16, chunk_vox a multiple
// of 32.  Work-groups (4 waves) run in XCD order: XCD c takes the
// contiguous range [c per, (c + 1) per), so the waves of one chunk and of
// the chunks above and beside it (whose rows its dz / dy taps read) share
// one L2 instead of every XCD fetching every chunk's rows.
template <int CT, int IT, int DXT, int DYT>
__global__ __launch_bounds__(256) void conv3d_wgrad_cf(const u16 *__restrict__ Xc, const u16 *__restrict__ Gc, int n,
                                                       int cin, int cout, int Kp, int chunk_vox, int cotg, int citg,
                                                       float *__restrict__ P, unsigned per) {
    static_assert(DYT == 1 || DXT == 3, "three y rows take the three x taps of each");
    constexpr int kGroups = 27 / (DXT * DYT);  // dz planes, (dz, dy) rows or single taps
    const int lane = threadIdx.x & 63;
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t chunks = (nvox + chunk_vox - 1) / chunk_vox;
    const int64_t tasks = (int64_t)kGroups * cotg * citg;
    const int64_t wave = (int64_t)lgdhip::xcd_tile(blockIdx.x, per, true) * 4 + threadIdx.x / 64;
    if (wave >= chunks * tasks) return;  // no barriers in this kernel
    const int64_t chunk = wave / tasks;
    int64_t rem = wave % tasks;
    const int grp = (int)(rem % kGroups);  // taps fastest: neighbouring waves share the G fragments
    rem /= kGroups;
    const int cig = (int)(rem % citg), cog = (int)(rem / citg);
    const int tap0 = grp * DXT * DYT;  // taps tap0 + 3 r + e: row dy + r, x offset e - 1 (DXT = 3)
    int dz, dy, dx0;
    tap_offset(tap0, 3, dz, dy, dx0);
    const int i = lane & 15, q = lane >> 4;
    f32x4 acc[DYT][DXT][CT][IT];
#pragma unroll
    for (int r = 0; r < DYT; ++r)
#pragma unroll
        for (int d = 0; d < DXT; ++d)
#pragma unroll
            for (int t = 0; t < CT; ++t)
#pragma unroll
                for (int u = 0; u < IT; ++u) acc[r][d][t][u] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int64_t v_begin = chunk * chunk_vox;
    const int64_t v_end = v_begin + chunk_vox < nvox ? v_begin + chunk_vox : nvox;
    // the fetch front's voxel v = vb + 8q of this lane group and its
    // coordinates, stepped along with v (32 voxels per step) instead of
    // divided out each time
    int x, y, z;
    {
        const int64_t v0 = v_begin + 8 * q;
        x = (int)(v0 % n);
        const int64_t r0 = v0 / n;
        y = (int)(r0 % n);
        z = (int)(r0 / n);
    }
    // One 32-voxel step's fragments: dY (A) and each X row's centre 8 with
    // the adjacent words of its left and right neighbours.  Loads outside
    // the volume read a zero block instead of being skipped, so they need
    // no branch.  Addresses are stepped, not recomputed: the voxel-blocked
    // copies put the front's 8-voxel group g at element (g C + channel) 8,
    // a step is 4 groups, tap (dz, dy) reads group g + (dz n^2 + dy n) / 8
    // (exact, n % 8 == 0) and the next row n cin elements further; the
    // coordinates only decide which loads are inside the volume.
    const gh_ptr zero = (gh_ptr)(const void *)kZero16;
    struct Step {
        uint4 a[CT], c[DYT][IT];
        unsigned prev[DYT][IT], next[DYT][IT];  // the neighbouring 8 voxels' adjacent word
    };
    const int64_t g0 = (v_begin + 8 * q) >> 3;
    const int64_t tap_groups = ((int64_t)dz * n * n + (int64_t)dy * n) >> 3;
    const int64_t row_elems = (int64_t)n * cin;
    gh_ptr pa[CT], px[IT];
#pragma unroll
    for (int t = 0; t < CT; ++t) pa[t] = (gh_ptr)Gc + ((g0 * cout + (cog * CT + t) * 16 + i) << 3);
#pragma unroll
    for (int u = 0; u < IT; ++u) px[u] = (gh_ptr)Xc + (((g0 + tap_groups) * cin + (cig * IT + u) * 16 + i) << 3);
    const int64_t step_a = (int64_t)32 * cout, step_x = (int64_t)32 * cin;  // 4 groups of 8 voxels
    auto fetch = [&](Step &f) {
#pragma unroll
        for (int t = 0; t < CT; ++t) f.a[t] = load_global16(pa[t]);
        const bool zok = (unsigned)(z + dz) < (unsigned)n;
        const bool xlo = (DXT == 3 || dx0 < 0) && x > 0;
        const bool xhi = (DXT == 3 || dx0 > 0) && x + 8 < n;
#pragma unroll
        for (int r = 0; r < DYT; ++r) {
            const bool rowok = zok && (unsigned)(y + dy + r) < (unsigned)n;
            const bool lo = rowok && xlo, hi = rowok && xhi;
#pragma unroll
            for (int u = 0; u < IT; ++u) {
                const gh_ptr pr = px[u] + r * row_elems;
                f.c[r][u] = load_global16(rowok ? pr : zero);
                f.prev[r][u] = load_global4(lo ? pr - 8 * cin + 6 : zero);
                f.next[r][u] = load_global4(hi ? pr + 8 * cin : zero);
            }
        }
    };
    // 32 voxels on: at most 32 / n <= 4 row wraps (n >= 8) and one plane
    // wrap (n^2 >= 64), as selects -- no data-dependent loop between a
    // step's loads and the MFMAs that wait for them
    auto advance = [n](int &xa, int &ya, int &za) {
        xa += 32;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bool w = xa >= n;
            xa = w ? xa - n : xa;
            ya = w ? ya + 1 : ya;
        }
        const bool wz = ya >= n;
        ya = wz ? ya - n : ya;
        za = wz ? za + 1 : za;
    };
    // Two step buffers, unrolled by two with fixed roles: step k + 1's
    // loads go out before step k's MFMAs, and step k + 2's before step
    // k + 1's, so a step's loads have a whole other step to arrive (a copy
    // from a "next" buffer would wait for them at the end of every step).
    // The fetch front stops at the chunk's last step, which later fetches
    // re-load; an odd step count ends with one step of a zero A fragment
    // (adding exact zeros), so the body has no branch.
    const int S = (int)((v_end - v_begin) / 32);
    int front = 0;  // step the next fetch loads
    auto fetch_front = [&](Step &f) {
        fetch(f);
        const bool on = front + 1 < S;
        int xa = x, ya = y, za = z;
        advance(xa, ya, za);
        front += on ? 1 : 0;
        x = on ? xa : x;
        y = on ? ya : y;
        z = on ? za : z;
#pragma unroll
        for (int t = 0; t < CT; ++t) pa[t] += on ? step_a : 0;
#pragma unroll
        for (int u = 0; u < IT; ++u) px[u] += on ? step_x : 0;
    };
    auto compute = [&](const Step &cur, bool live) {
        Frag a[CT];
#pragma unroll
        for (int t = 0; t < CT; ++t) a[t].q = live ? cur.a[t] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int r = 0; r < DYT; ++r) {
            Frag b[DXT][IT];
#pragma unroll
            for (int u = 0; u < IT; ++u) {
                if (DXT == 3) {
                    b[0][u].q = shift_from_prev(cur.c[r][u], cur.prev[r][u]);
                    b[1 % DXT][u].q = cur.c[r][u];
                    b[2 % DXT][u].q = shift_from_next(cur.c[r][u], cur.next[r][u]);
                } else {
                    b[0][u].q = dx0 == 0 ? cur.c[r][u]
                                         : (dx0 < 0 ? shift_from_prev(cur.c[r][u], cur.prev[r][u])
                                                    : shift_from_next(cur.c[r][u], cur.next[r][u]));
                }
            }
#pragma unroll
            for (int e = 0; e < DXT; ++e)
#pragma unroll
                for (int t = 0; t < CT; ++t)
#pragma unroll
                    for (int u = 0; u < IT; ++u)
                        acc[r][e][t][u] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t].v, b[e][u].v, acc[r][e][t][u], 0, 0, 0);
        }
    };
    Step s0, s1;
    fetch_front(s0);
    for (int k = 0; k < S; k += 2) {
        fetch_front(s1);  // step k + 1
        compute(s0, true);
        fetch_front(s0);  // step k + 2
        compute(s1, k + 1 < S);
    }
    // D lane layout: row (co) 4*(lane/16) + e, column (ci) lane % 16
#pragma unroll
    for (int r = 0; r < DYT; ++r)
#pragma unroll
        for (int d = 0; d < DXT; ++d)
#pragma unroll
            for (int t = 0; t < CT; ++t)
#pragma unroll
                for (int u = 0; u < IT; ++u)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int co = (cog * CT + t) * 16 + 4 * q + e;
                        const int ci = (cig * IT + u) * 16 + i;
                        P[((int64_t)chunk * cout + co) * Kp + (tap0 + 3 * r + d) * cin + ci] = acc[r][d][t][u][e];
                    }
}

// dW[co][k] = sum over chunks of P (fixed order) for k < K; the padding
// columns k >= K stay zero
// A work-group of 64 x 4 threads takes 64 consecutive outputs; thread row j
// sums the chunks j, j + 4, j + 8, ... (eight loads in flight, coalesced
// across the 64 outputs) and the four row sums are added in row order:
// a fixed order for every output.
constexpr int kWgradReduceOutputs = 64, kWgradReduceRows = 4;
__global__ __launch_bounds__(256) void conv3d_wgrad_reduce(const float *__restrict__ P, int chunks, int cout, int Kp,
                                                           int K, float *__restrict__ dW, int accumulate) {
    __shared__ float part[kWgradReduceRows][kWgradReduceOutputs];
    const int total = cout * Kp;  // < 2^31: cout <= 192, Kp <= 27 * 192 + 31
    const int o = (int)blockIdx.x * kWgradReduceOutputs + (int)threadIdx.x;
    const int j = (int)threadIdx.y;
    float s = 0.0f;
    if (o < total && o % Kp < K) {
#pragma unroll 8
        for (int c = j; c < chunks; c += kWgradReduceRows) s += P[(int64_t)c * total + o];
    }
    part[j][threadIdx.x] = s;
    __syncthreads();
    if (j != 0 || o >= total) return;
    float r = part[0][threadIdx.x];
#pragma unroll
    for (int k = 1; k < kWgradReduceRows; ++k) r += part[k][threadIdx.x];
    if (o % Kp >= K) {  // padding columns stay zero
        dW[o] = accumulate ? dW[o] : 0.0f;
        return;
    }
    dW[o] = accumulate ? dW[o] + r : r;
}

// db from the voxel-blocked gradient (C channels): partial sums of each
// channel over chunks of 8192 voxels, then a fixed-order sum of the chunks
constexpr int kBiasChunk = 8192;
__global__ __launch_bounds__(256) void row_sum_partial(const u16 *__restrict__ Gc, int64_t nvox, float *__restrict__ Pb,
                                                       int bchunks, int C) {
    __shared__ float red[256];
    const int co = blockIdx.x, c = blockIdx.y;
    const int64_t v0 = (int64_t)c * kBiasChunk;
    const int64_t v1 = v0 + kBiasChunk < nvox ? v0 + kBiasChunk : nvox;
    float s = 0.0f;
    for (int64_t v = v0 + threadIdx.x; v < v1; v += 256) s += bf2f(Gc[((((v >> 3) * C) + co) << 3) + (v & 7)]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) Pb[co * bchunks + c] = red[0];
}
__global__ void row_sum_final(const float *__restrict__ Pb, int bchunks, int cout, float *__restrict__ db,
                              int accumulate) {
    const int co = blockIdx.x * blockDim.x + threadIdx.x;
    if (co >= cout) return;
    float s = 0.0f;
    for (int c = 0; c < bchunks; ++c) s += Pb[co * bchunks + c];
    db[co] = accumulate ? db[co] + s : s;
}

// db[co] partial sums over chunks of 8192 voxels, channels-last input:
// thread t of a block handles channel t % cout of every (256 / cout)-th
// voxel, so a wave reads whole voxel rows (coalesced); fixed-order sums
__global__ __launch_bounds__(256) void bias_grad_partial(const u16 *__restrict__ dY, int64_t nvox, int ldy, int offy,
                                                         int cout, float *__restrict__ Pb, int bchunks) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    const int lanes = 256 / cout;  // voxel lanes
    const int t = threadIdx.x;
    float s = 0.0f;
    if (t < lanes * cout) {
        const int co = t % cout, vl = t / cout;
        const int64_t v0 = (int64_t)c * kBiasChunk;
        const int64_t v1 = v0 + kBiasChunk < nvox ? v0 + kBiasChunk : nvox;
        for (int64_t v = v0 + vl; v < v1; v += lanes) s += bf2f(dY[v * ldy + offy + co]);
    }
    red[t] = s;
    __syncthreads();
    if (t < cout) {
        float sum = 0.0f;
        for (int vl = 0; vl < lanes; ++vl) sum += red[vl * cout + t];
        Pb[t * bchunks + c] = sum;
    }
}

// dgrad weights: Wt[ci][tap'*cout + co] = W[co][tap*cin + ci], tap' = 26 - tap
__global__ void conv3d_dgrad_weights(const u16 *__restrict__ W, int cin, int cout, int ksize, int Kp,
                                     u16 *__restrict__ Wt, int Kpt) {
    const int taps = ksize == 1 ? 1 : 27;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)cin * Kpt) return;
    const int ci = (int)(idx / Kpt);
    const int kk = (int)(idx % Kpt);
    u16 v = 0;
    if (kk < taps * cout) {
        const int tp = kk / cout, co = kk % cout;
        const int tap = taps - 1 - tp;
        v = W[(int64_t)co * Kp + tap * cin + ci];
    }
    Wt[idx] = v;
}

// g *= (y > 0) for the ReLU backward (in place, BF16)
__global__ void relu_backward(u16 *g, int ldg, int offg, const u16 *__restrict__ y, int ldy, int offy, int c,
                              int64_t nvox) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nvox * c) return;
    const int64_t v = idx / c;
    const int ch = (int)(idx % c);
    if (!(bf2f(y[v * ldy + offy + ch]) > 0.0f)) g[v * ldg + offg + ch] = 0;
}

// 2x2x2 average pool: Y (n/2)^3 <- X n^3
__global__ void avgpool2(const u16 *__restrict__ X, int n, int ldx, int offx, u16 *Y, int ldy, int offy, int c) {
    const int m = n / 2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * m * m * c) return;
    const int ch = (int)(idx % c);
    const int64_t v = idx / c;
    const int x = (int)(v % m), y = (int)((v / m) % m), z = (int)(v / ((int64_t)m * m));
    float s = 0.0f;
    for (int d = 0; d < 8; ++d) {
        const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
        s += bf2f(X[((int64_t)((zz * n + yy) * n + xx)) * ldx + offx + ch]);
    }
    Y[v * ldy + offy + ch] = f2bf(s * 0.125f);
}

// gradient of the average pool: dX[child] (+)= dY[parent] / 8
__global__ void avgpool2_backward(const u16 *__restrict__ dY, int n, int ldy, int offy, u16 *dX, int ldx, int offx,
                                  int c, int accumulate) {
    const int m = n / 2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n * n * c) return;
    const int ch = (int)(idx % c);
    const int64_t v = idx / c;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    const int64_t p = ((int64_t)(z / 2) * m + y / 2) * m + x / 2;
    float g = bf2f(dY[p * ldy + offy + ch]) * 0.125f;
    u16 *q = dX + v * ldx + offx + ch;
    if (accumulate) g += bf2f(*q);
    *q = f2bf(g);
}

// nearest 2x upsample: Y (2m)^3 <- X m^3
__global__ void upsample2(const u16 *__restrict__ X, int m, int ldx, int offx, u16 *Y, int ldy, int offy, int c) {
    const int n = 2 * m;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n * n * c) return;
    const int ch = (int)(idx % c);
    const int64_t v = idx / c;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    const int64_t p = ((int64_t)(z / 2) * m + y / 2) * m + x / 2;
    Y[v * ldy + offy + ch] = X[p * ldx + offx + ch];
}

// gradient of the upsample: dX[parent] = sum of the 8 children
__global__ void upsample2_backward(const u16 *__restrict__ dY, int m, int ldy, int offy, u16 *dX, int ldx, int offx,
                                   int c, int accumulate) {
    const int n = 2 * m;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * m * m * c) return;
    const int ch = (int)(idx % c);
    const int64_t v = idx / c;
    const int x = (int)(v % m), y = (int)((v / m) % m), z = (int)(v / ((int64_t)m * m));
    float s = 0.0f;
    for (int d = 0; d < 8; ++d) {
        const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
        s += bf2f(dY[((int64_t)((zz * n + yy) * n + xx)) * ldy + offy + ch]);
    }
    u16 *q = dX + v * ldx + offx + ch;
    if (accumulate) s += bf2f(*q);
    *q = f2bf(s);
}

// The same four resampling kernels over 8 channels per thread (16-byte loads
// and stores), used when the channel count, row strides, channel offsets and
// base pointers are multiples of 8 elements.  Each channel is summed in the
// order of the kernels above, so the results are bit-identical.
struct alignas(16) Bf8 {
    u16 h[8];
};
__device__ __forceinline__ Bf8 ld8(const u16 *p) { return *reinterpret_cast<const Bf8 *>(p); }
__device__ __forceinline__ void st8(u16 *p, const Bf8 &v) { *reinterpret_cast<Bf8 *>(p) = v; }

// Thread coordinates of the 8-channel resampling kernels over a w^3 volume:
// a launch row_grid(w, c8) gives each (y, z) row blockIdx.y / .z, and the
// threads along x cover its (x, channel chunk) pairs, the chunk fastest
// (consecutive threads, consecutive 16-byte pieces of a voxel's channels).
// 32-bit arithmetic throughout (a flat 64-bit index divided by run-time
// extents costs hundreds of instructions per thread).
__device__ __forceinline__ bool row_thread(int w, int c8, int &x, int &ch) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= w * c8) return false;
    x = t / c8;
    ch = (t - x * c8) * 8;
    return true;
}

__global__ void avgpool2_v8(const u16 *__restrict__ X, int n, int ldx, int offx, u16 *Y, int ldy, int offy, int c8) {
    const int m = n / 2;
    int x, ch;
    if (!row_thread(m, c8, x, ch)) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const int64_t v = ((int64_t)z * m + y) * m + x;
    float s[8] = {};
    for (int d = 0; d < 8; ++d) {
        const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
        const Bf8 a = ld8(X + ((int64_t)((zz * n + yy) * n + xx)) * ldx + offx + ch);
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += bf2f(a.h[k]);
    }
    Bf8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o.h[k] = f2bf(s[k] * 0.125f);
    st8(Y + v * ldy + offy + ch, o);
}

__global__ void avgpool2_backward_v8(const u16 *__restrict__ dY, int n, int ldy, int offy, u16 *dX, int ldx, int offx,
                                     int c8, int accumulate) {
    const int m = n / 2;
    int x, ch;
    if (!row_thread(n, c8, x, ch)) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const int64_t v = ((int64_t)z * n + y) * n + x;
    const int64_t p = ((int64_t)(z / 2) * m + y / 2) * m + x / 2;
    const Bf8 a = ld8(dY + p * ldy + offy + ch);
    u16 *q = dX + v * ldx + offx + ch;
    Bf8 old;
    if (accumulate) old = ld8(q);
    Bf8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float g = bf2f(a.h[k]) * 0.125f;
        if (accumulate) g += bf2f(old.h[k]);
        o.h[k] = f2bf(g);
    }
    st8(q, o);
}

__global__ void upsample2_v8(const u16 *__restrict__ X, int m, int ldx, int offx, u16 *Y, int ldy, int offy, int c8) {
    const int n = 2 * m;
    int x, ch;
    if (!row_thread(n, c8, x, ch)) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const int64_t v = ((int64_t)z * n + y) * n + x;
    const int64_t p = ((int64_t)(z / 2) * m + y / 2) * m + x / 2;
    st8(Y + v * ldy + offy + ch, ld8(X + p * ldx + offx + ch));
}

// Nearest 2x upsampling of x (m^3, cu channels) written together with a
// skip tensor (n^3 = (2m)^3, cs channels) as whole concatenation rows:
//   y[v][offy + c] = c < cu ? x[parent(v)][offx + c] : s[v][offs + c - cu]
// so every 128-byte line of y is written by one launch in full (the
// encoder writing its channels into the rows and the upsampling writing the
// others leave each line partly written twice, and a partial line costs
// the memory a read-modify-write).  8 channels per thread; consecutive
// threads write consecutive 16-byte pieces of a row.
__global__ void upsample2_concat_v8(const u16 *__restrict__ X, int m, int ldx, int offx, int cu8,
                                    const u16 *__restrict__ S, int lds, int offs, u16 *Y, int ldy, int offy, int c8) {
    const int n = 2 * m;
    int x, ch;
    if (!row_thread(n, c8, x, ch)) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const int64_t v = ((int64_t)z * n + y) * n + x;
    Bf8 val;
    if (ch < 8 * cu8) {
        const int64_t p = ((int64_t)(z / 2) * m + y / 2) * m + x / 2;
        val = ld8(X + p * ldx + offx + ch);
    } else {
        val = ld8(S + v * lds + offs + ch - 8 * cu8);
    }
    st8(Y + v * ldy + offy + ch, val);
}

// the same, one element per thread (any channel counts, strides, offsets)
__global__ void upsample2_concat(const u16 *__restrict__ X, int m, int ldx, int offx, int cu,
                                 const u16 *__restrict__ S, int lds, int offs, int cs, u16 *Y, int ldy, int offy) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = cu + cs, n = 2 * m;
    if (idx >= (int64_t)n * n * n * c) return;
    const int ch = (int)(idx % c);
    const int64_t v = idx / c;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    u16 val;
    if (ch < cu) {
        const int64_t p = ((int64_t)(z / 2) * m + y / 2) * m + x / 2;
        val = X[p * ldx + offx + ch];
    } else {
        val = S[v * lds + offs + ch - cu];
    }
    Y[v * ldy + offy + ch] = val;
}

__global__ void upsample2_backward_v8(const u16 *__restrict__ dY, int m, int ldy, int offy, u16 *dX, int ldx, int offx,
                                      int c8, int accumulate) {
    const int n = 2 * m;
    int x, ch;
    if (!row_thread(m, c8, x, ch)) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const int64_t v = ((int64_t)z * m + y) * m + x;
    float s[8] = {};
    for (int d = 0; d < 8; ++d) {
        const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
        const Bf8 a = ld8(dY + ((int64_t)((zz * n + yy) * n + xx)) * ldy + offy + ch);
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += bf2f(a.h[k]);
    }
    u16 *q = dX + v * ldx + offx + ch;
    Bf8 old;
    if (accumulate) old = ld8(q);
    Bf8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o.h[k] = f2bf(accumulate ? s[k] + bf2f(old.h[k]) : s[k]);
    st8(q, o);
}

// FP32 -> BF16 into a channel slot of a channels-last tensor
__global__ void pack_bf16(const float *__restrict__ src, int64_t nvox, u16 *Y, int ldy, int offy) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nvox) Y[v * ldy + offy] = f2bf(src[v]);
}

// MSE loss: L = mean((pred - target)^2); grad = 2 (pred - target) / N as BF16;
// per-block partial sums of the squared error (deterministic reduction)
__global__ void mse_loss(const float *__restrict__ pred, const float *__restrict__ target, int64_t nvox, u16 *grad,
                         int ldg, float *__restrict__ partial) {
    __shared__ float lds[256];
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float se = 0.0f;
    if (v < nvox) {
        const float d = pred[v] - target[v];
        se = d * d;
        grad[v * ldg] = f2bf(2.0f * d / (float)nvox);
    }
    lds[threadIdx.x] = se;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

// Adam on FP32 master weights [cout][Kp] (padding columns stay zero); writes BF16 copies
__global__ void adam_step(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                          const float *__restrict__ g, int64_t count, int cols_used, int Kp, float lr, float b1,
                          float b2, float eps, float c1, float c2, u16 *__restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (Kp > 0 && (int)(i % Kp) >= cols_used) {
        w[i] = 0.0f;
        if (wb) wb[i] = 0;
        return;
    }
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }
// the launch of a row_thread kernel over a w^3 volume with c8 chunks of 8
// channels (w <= kMaxRowGrid: the grid's y and z extents; larger volumes take
// the per-element kernels)
constexpr int64_t kMaxRowGrid = 65535;
dim3 row_grid(int w, int c8) { return dim3(blocks_for((int64_t)w * c8), (unsigned)w, (unsigned)w); }
// 8-channel (16-byte) resampling applies: channel count, strides and channel
// offsets multiples of 8 elements, both base pointers 16-byte aligned
bool vec8(const void *a, const void *b, int c, int lda, int offa, int ldb, int offb) {
    return c % 8 == 0 && lda % 8 == 0 && offa % 8 == 0 && ldb % 8 == 0 && offb % 8 == 0 &&
           ((uintptr_t)a & 15) == 0 && ((uintptr_t)b & 15) == 0;
}

// shapes the voxel-blocked weight gradient (conv3d_wgrad_cf) handles
bool wgrad_cf_shape(int n, int cin, int ksize, int cout) {
    return ksize == 3 && cin % 16 == 0 && cout % 16 == 0 && n % 8 == 0;
}

// conv3d_wgrad_cf's tiling: CT x IT tiles of 16 co x 16 ci per wave, the
// three x taps of a row at once (DXT = 3) while the accumulators stay
// within 48 registers, and for a single tile the three rows of a dz too
// (DYT = 3, 36 accumulator registers)
struct CfTiling {
    int CT, IT, DXT, DYT, cotg, citg;
    int tasks() const { return 27 / (DXT * DYT) * cotg * citg; }  // waves per chunk
};
CfTiling cf_tiling(int cin, int cout) {
    const int cot = cout / 16, cit = cin / 16;
    CfTiling t;
    t.CT = cot % 4 == 0 ? 4 : (cot % 2 == 0 ? 2 : 1);
    t.IT = cit % 4 == 0 ? 4 : (cit % 2 == 0 ? 2 : 1);
    t.cotg = cot / t.CT;
    t.citg = cit / t.IT;
    t.DXT = t.CT * t.IT <= 4 ? 3 : 1;
    t.DYT = t.CT * t.IT == 1 ? 3 : 1;  // one co and one ci tile: all three rows of a dz
    return t;
}

// Voxels per chunk when the caller leaves it to the library (chunk_vox <=
// 0): the largest power of two <= 4096 (and >= 128) that still gives the
// launch 2048 waves (8 per CU), since each wave marches its chunk 32 voxels
// per dependent step.  4096 at 128^3 (16 channels: 4608 waves of 128
// steps); the 32^3 and 16^3 levels would otherwise run 54-162 work-groups.
int wgrad_chunk(int n, int cin, int ksize, int cout) {
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t tasks = wgrad_cf_shape(n, cin, ksize, cout) ? cf_tiling(cin, cout).tasks()
                                                               : lgdhip_conv3d_kpad(cin, ksize) / 16;
    int c = 4096;
    while (c > 128 && (nvox + c - 1) / c * tasks < 2048) c /= 2;
    return c;
}

int check_tensor(int n, int ld, int off, int c) {
    if (n < 1 || ld < 1 || off < 0 || c < 1 || off + c > ld)
        return lgdhip::set_error(LGDHIP_ERR_INVALID_ARG, "bad tensor shape (n, ld, off, c)");
    if ((int64_t)n * n * n * ld >= ((int64_t)1 << 40))
        return lgdhip::set_error(LGDHIP_ERR_UNSUPPORTED, "tensor too large");
    return 0;
}

} // namespace

extern "C" {

int lgdhip_conv3d_kpad(int cin, int ksize) { return (((ksize == 1 ? 1 : 27) * cin + 31) / 32) * 32; }

size_t lgdhip_conv3d_forward_workspace_bytes(int n, int cin, int ksize, int cout) {
    if (n < 1 || cin < 1 || cout < 1 || !gemm_applies(n, cin, ksize, cout)) return 0;
    return gemm_workspace_bytes(n, gemm_plan(n, cin, cout, true));
}

int lgdhip_conv3d_bf16_forward(const void *x, int n, int ldx, int offx, int cin, int ksize, const void *w,
                               const float *bias, int cout, void *y, int ldy, int offy, float *y32, int ld32,
                               int relu, int accumulate, lgdhip_stream_t stream) {
    return lgdhip_conv3d_bf16_forward_ws(x, n, ldx, offx, cin, ksize, w, bias, cout, y, ldy, offy, y32, ld32, relu,
                                         accumulate, nullptr, 0, stream);
}

int lgdhip_conv3d_bf16_forward_ws(const void *x, int n, int ldx, int offx, int cin, int ksize, const void *w,
                                  const float *bias, int cout, void *y, int ldy, int offy, float *y32, int ld32,
                                  int relu, int accumulate, void *workspace, size_t workspace_bytes,
                                  lgdhip_stream_t stream) {
    LGD_REQUIRE(x && w && (y || y32), "NULL pointer");
    LGD_REQUIRE(ksize == 1 || ksize == 3, "ksize must be 1 or 3");
    LGD_REQUIRE(cout >= 1 && cout <= 16 * kMaxNTiles, "cout must be in [1, 192]");
    LGD_REQUIRE(!y32 || ld32 >= cout, "ld32 must be >= cout");
    int rc = check_tensor(n, ldx, offx, cin);
    if (rc) return rc;
    if (y) {
        rc = check_tensor(n, ldy, offy, cout);
        if (rc) return rc;
    }
    if ((cin & 7) == 0) {
        LGD_REQUIRE((ldx & 7) == 0 && (offx & 7) == 0, "ldx and offx must be multiples of 8 when cin % 8 == 0");
        LGD_REQUIRE(((uintptr_t)x & 15) == 0, "x must be 16-byte aligned");
    }
    LGD_REQUIRE(((uintptr_t)w & 15) == 0, "w must be 16-byte aligned");
    const int Kp = lgdhip_conv3d_kpad(cin, ksize);
    const int nt = (cout + 15) / 16;
    hipStream_t s = lgdhip::as_stream(stream);
    const u16 *X = (const u16 *)x;
    const u16 *Wb = (const u16 *)w;
    u16 *Y = (u16 *)y;
    const int64_t nvox = (int64_t)n * n * n;
    if (ksize == 1 && cout == 1 && (cin & 7) == 0) {  // ldx, offx, alignment checked above
        conv1x1_to1<<<(unsigned)((nvox + 255) / 256), 256, 0, s>>>(X, nvox, ldx, offx, cin, Wb, bias, Y, ldy, offy,
                                                                    y32, ld32, relu, accumulate);
        LGD_LAUNCHED("conv1x1_to1");
        return 0;
    }
    // z-marching kernel: 16 output channels written as plain BF16 into an
    // aligned slice, input of at most 6 groups of 8 channels (or a masked
    // row of fewer than 8).  Wider outputs on volumes of at least
    // kZmSliceMinN^3 run one launch per 16-channel slice (weights rows 16 j
    // .. 16 j + 15, output channels offy + 16 j ..): each launch still has
    // hundreds of work-groups, each input plane is staged once per slice,
    // and the slices together beat one GEMM-tiled launch (the input
    // gradient of a 16 -> 48 layer at 128^3: 412 us there)
    const bool zm_width = cout == 16 || (cout % 16 == 0 && cout <= 16 * kMaxNTiles && n >= kZmSliceMinN);
    if (ksize == 3 && zm_width && n >= kLdsMinN && y && !y32 && !accumulate && (ldy & 7) == 0 &&
        (offy & 7) == 0 && ((uintptr_t)y & 15) == 0) {
        const bool masked = cin < 8 && (ldx & 7) == 0 && (offx & 7) == 0 && offx + 8 <= ldx &&
                            ((uintptr_t)x & 15) == 0;
        const int groups = (cin & 7) == 0 ? cin / 8 : (masked ? 1 : 0);
        if (groups >= 1 && groups <= 6) {
            const int bx = (n + 15) / 16, by = (n + 7) / 8, bz = (n + kZmChunk - 1) / kZmChunk;
            const unsigned blocks = (unsigned)((int64_t)bx * by * bz);
            for (int j = 0; j < cout / 16; ++j) {
                const u16 *Wj = Wb + (size_t)16 * j * Kp;  // Kp % 32 == 0: 16-byte aligned
                const float *bj = bias ? bias + 16 * j : nullptr;
                const int offj = offy + 16 * j;
#define LGD_ZM_CASE(G, I)                                                                                    \
    if (groups == G && (I == kInMasked) == masked)                                                           \
        conv3d_fwd_zm<G, I><<<blocks, 256, 0, s>>>(X, n, ldx, offx, cin, Wj, Kp, bj, 16, Y, ldy, offj, relu, bx, by);
                LGD_ZM_CASE(1, kInMasked) LGD_ZM_CASE(1, kInVec) LGD_ZM_CASE(2, kInVec) LGD_ZM_CASE(3, kInVec)
                LGD_ZM_CASE(4, kInVec) LGD_ZM_CASE(5, kInVec) LGD_ZM_CASE(6, kInVec)
#undef LGD_ZM_CASE
                LGD_LAUNCHED("conv3d_fwd_zm");
            }
            return 0;
        }
    }
    if (gemm_applies(n, cin, ksize, cout)) {
        const GemmPlan p = gemm_plan(n, cin, cout, workspace != nullptr);
        if (workspace && workspace_bytes < gemm_workspace_bytes(n, p))
            return lgdhip::set_error(LGDHIP_ERR_WORKSPACE,
                                     "forward workspace too small (lgdhip_conv3d_forward_workspace_bytes)");
        const unsigned blocks = (unsigned)((int64_t)p.mblocks * p.nblocks * p.slices);
        if (p.slices > 1) {
            // (gemm_plan splits only launches of fewer than 4 x 256 blocks)
            float *P = (float *)workspace;
            conv3d_fwd_gemm<true><<<blocks, 256, 0, s>>>(X, n, ldx, offx, cin, Wb, Kp, bias, cout, Y, ldy, offy, y32,
                                                         ld32, relu, accumulate, p.mblocks, p.nblocks, p.slices,
                                                         p.stages, P);
            LGD_LAUNCHED("conv3d_fwd_gemm");
            conv3d_gemm_reduce<<<blocks_for(nvox * cout), 256, 0, s>>>(P, nvox, p.nblocks * kGemmBN, p.slices, bias,
                                                                       cout, Y, ldy, offy, y32, ld32, relu,
                                                                       accumulate);
            LGD_LAUNCHED("conv3d_gemm_reduce");
        } else {
            conv3d_fwd_gemm<false><<<blocks, 256, 0, s>>>(X, n, ldx, offx, cin, Wb, Kp, bias, cout, Y, ldy, offy, y32,
                                                          ld32, relu, accumulate, p.mblocks, p.nblocks, 1, p.stages,
                                                          nullptr);
            LGD_LAUNCHED("conv3d_fwd_gemm");
        }
        return 0;
    }
    if (ksize == 3 && nt <= 2 && n >= kLdsMinN) {
        const int bx = (n + kLdsTX - 1) / kLdsTX, by = (n + kLdsTY - 1) / kLdsTY;
        const int bz = (n + 8 / nt - 1) / (8 / nt);
        const unsigned blocks = (unsigned)((int64_t)bx * by * bz);
        // input: 16-byte loads when cin % 8 == 0 (ldx, offx and the alignment
        // were checked above), or for cin < 8 when the 8-channel row
        // [offx, offx + 8) lies inside the voxel's ldx channels and is aligned
        const bool aligned_row = cin < 8 && (ldx & 7) == 0 && (offx & 7) == 0 && offx + 8 <= ldx &&
                                 ((uintptr_t)x & 15) == 0;
        const int in = (cin & 7) == 0 ? kInVec : (aligned_row ? kInMasked : kInElem);
        // output: 16-byte stores of full cout tiles into an aligned slice
        const bool vst = y && !y32 && !accumulate && cout == 16 * nt && (ldy & 7) == 0 && (offy & 7) == 0 &&
                         ((uintptr_t)y & 15) == 0;
#define LGD_LDS_CASE(T, I, V)                                                                                   \
    if (nt == T && in == I && vst == V)                                                                         \
        conv3d_fwd_lds<T, I, V><<<blocks, 256, 0, s>>>(X, n, ldx, offx, cin, Wb, Kp, bias, cout, Y, ldy, offy, y32, \
                                                       ld32, relu, accumulate, bx, by);
        LGD_LDS_CASE(1, kInVec, true) LGD_LDS_CASE(1, kInVec, false) LGD_LDS_CASE(1, kInMasked, true)
        LGD_LDS_CASE(1, kInMasked, false) LGD_LDS_CASE(1, kInElem, true) LGD_LDS_CASE(1, kInElem, false)
        LGD_LDS_CASE(2, kInVec, true) LGD_LDS_CASE(2, kInVec, false) LGD_LDS_CASE(2, kInMasked, true)
        LGD_LDS_CASE(2, kInMasked, false) LGD_LDS_CASE(2, kInElem, true) LGD_LDS_CASE(2, kInElem, false)
#undef LGD_LDS_CASE
        LGD_LAUNCHED("conv3d_fwd_lds");
        return 0;
    }
    const FwdShape sh = fwd_shape(nvox, nt);
#define LGD_CONV_CASE(T, M)                                                                                   \
    if (sh.nt == T && sh.mt == M) {                                                                           \
        const int64_t waves = (nvox + 16 * M - 1) / (16 * M) * sh.groups;                                     \
        conv3d_fwd_mfma<T, M><<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(                                    \
            X, n, ldx, offx, cin, ksize, Wb, Kp, bias, cout, Y, ldy, offy, y32, ld32, relu, accumulate,        \
            sh.groups);                                                                                       \
    }
    LGD_CONV_CASE(1, 4) LGD_CONV_CASE(1, 2) LGD_CONV_CASE(1, 1)
    LGD_CONV_CASE(2, 4) LGD_CONV_CASE(2, 2) LGD_CONV_CASE(2, 1)
    LGD_CONV_CASE(3, 4) LGD_CONV_CASE(3, 2) LGD_CONV_CASE(3, 1)
    LGD_CONV_CASE(4, 4) LGD_CONV_CASE(4, 2) LGD_CONV_CASE(4, 1)
    LGD_CONV_CASE(5, 2) LGD_CONV_CASE(5, 1)
    LGD_CONV_CASE(6, 2) LGD_CONV_CASE(6, 1)
    LGD_CONV_CASE(7, 2) LGD_CONV_CASE(7, 1)
    LGD_CONV_CASE(8, 2) LGD_CONV_CASE(8, 1)
    LGD_CONV_CASE(9, 1)
    LGD_CONV_CASE(10, 1)
    LGD_CONV_CASE(11, 1)
    LGD_CONV_CASE(12, 1)
#undef LGD_CONV_CASE
    LGD_LAUNCHED("conv3d_fwd_mfma");
    return 0;
}

size_t lgdhip_conv3d_wgrad_workspace_bytes(int n, int cin, int ksize, int cout, int chunk_vox) {
    if (n < 1 || cin < 1 || cout < 1 || (ksize != 1 && ksize != 3)) return 0;
    if (chunk_vox <= 0) chunk_vox = wgrad_chunk(n, cin, ksize, cout);
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t chunks = (nvox + chunk_vox - 1) / chunk_vox;
    // [P: chunk partials][bias partials][x and dy voxel-blocked copies]
    size_t bytes = (size_t)(chunks * cout * lgdhip_conv3d_kpad(cin, ksize) * 4);
    bytes += (size_t)cout * (size_t)((nvox + kBiasChunk - 1) / kBiasChunk) * 4;
    if (wgrad_cf_shape(n, cin, ksize, cout)) bytes += 256 + (size_t)(cin + cout) * (size_t)nvox * 2;
    return bytes;
}

int lgdhip_conv3d_bf16_wgrad(const void *x, int n, int ldx, int offx, int cin, int ksize, const void *dy, int ldy,
                             int offy, int cout, float *dw, float *db, int accumulate, int chunk_vox,
                             void *workspace, size_t workspace_bytes, lgdhip_stream_t stream) {
    LGD_REQUIRE(x && dy && dw && workspace, "NULL pointer");
    LGD_REQUIRE(ksize == 1 || ksize == 3, "ksize must be 1 or 3");
    LGD_REQUIRE(cout >= 1 && cout <= 16 * kMaxNTiles, "cout must be in [1, 192]");
    int rc = check_tensor(n, ldx, offx, cin);
    if (rc) return rc;
    rc = check_tensor(n, ldy, offy, cout);
    if (rc) return rc;
    if (chunk_vox <= 0) chunk_vox = wgrad_chunk(n, cin, ksize, cout);
    LGD_REQUIRE(chunk_vox % 32 == 0, "chunk_vox must be a multiple of 32");
    if (workspace_bytes < lgdhip_conv3d_wgrad_workspace_bytes(n, cin, ksize, cout, chunk_vox))
        return lgdhip::set_error(LGDHIP_ERR_WORKSPACE, "wgrad workspace too small");
    const int Kp = lgdhip_conv3d_kpad(cin, ksize);
    const int K = (ksize == 1 ? 1 : 27) * cin;
    const int64_t nvox = (int64_t)n * n * n;
    const int chunks = (int)((nvox + chunk_vox - 1) / chunk_vox);
    hipStream_t s = lgdhip::as_stream(stream);
    float *P = (float *)workspace;
    const bool cf = wgrad_cf_shape(n, cin, ksize, cout) && ldx % 8 == 0 && offx % 8 == 0 && ldy % 8 == 0 &&
                    offy % 8 == 0 && (uintptr_t)x % 16 == 0 && (uintptr_t)dy % 16 == 0;
    if (cf) {
        // voxel-blocked copies: every MFMA fragment becomes one 16-byte load per lane
        const int bchunks0 = (int)((nvox + kBiasChunk - 1) / kBiasChunk);
        float *Pb = P + (size_t)chunks * cout * Kp;
        char *base = (char *)(Pb + (size_t)cout * bchunks0);
        base += (256 - (uintptr_t)base % 256) % 256;
        u16 *Xc = (u16 *)base;
        u16 *Gc = Xc + (size_t)cin * nvox;
        to_voxel_blocked<<<dim3(blocks_for(nvox), cin / 8), 256, 0, s>>>((const u16 *)x, nvox, ldx, offx, cin, Xc);
        to_voxel_blocked<<<dim3(blocks_for(nvox), cout / 8), 256, 0, s>>>((const u16 *)dy, nvox, ldy, offy, cout, Gc);
        LGD_LAUNCHED("to_voxel_blocked");
        const CfTiling tl = cf_tiling(cin, cout);
        const int CT = tl.CT, IT = tl.IT, DXT = tl.DXT, DYT = tl.DYT, cotg = tl.cotg, citg = tl.citg;
        const int64_t waves = (int64_t)chunks * tl.tasks();
        const unsigned wgs = (unsigned)((waves + 3) / 4);
        const unsigned per = lgdhip::xcd_per(wgs), grid = lgdhip::xcd_groups(wgs);
#define LGD_CF_CASE(A, B, D, R)                                                                              \
    if (CT == A && IT == B && DXT == D && DYT == R)                                                          \
        conv3d_wgrad_cf<A, B, D, R><<<grid, 256, 0, s>>>(Xc, Gc, n, cin, cout, Kp, chunk_vox, cotg, citg, P, per);
        LGD_CF_CASE(1, 1, 3, 3)
        LGD_CF_CASE(1, 2, 3, 1)
        LGD_CF_CASE(1, 4, 3, 1)
        LGD_CF_CASE(2, 1, 3, 1)
        LGD_CF_CASE(2, 2, 3, 1)
        LGD_CF_CASE(2, 4, 1, 1)
        LGD_CF_CASE(4, 1, 3, 1)
        LGD_CF_CASE(4, 2, 1, 1)
        LGD_CF_CASE(4, 4, 1, 1)
#undef LGD_CF_CASE
        LGD_LAUNCHED("conv3d_wgrad_cf");
        conv3d_wgrad_reduce<<<(unsigned)((cout * Kp + kWgradReduceOutputs - 1) / kWgradReduceOutputs),
                              dim3(kWgradReduceOutputs, kWgradReduceRows), 0, s>>>(P, chunks, cout, Kp, K, dw,
                                                                                    accumulate);
        LGD_LAUNCHED("conv3d_wgrad_reduce");
        if (db) {
            const int bchunks = (int)((nvox + kBiasChunk - 1) / kBiasChunk);
            row_sum_partial<<<dim3(cout, bchunks), 256, 0, s>>>(Gc, nvox, Pb, bchunks, cout);
            row_sum_final<<<blocks_for(cout), 256, 0, s>>>(Pb, bchunks, cout, db, accumulate);
            LGD_LAUNCHED("row_sum");
        }
        return 0;
    }
    const int64_t waves = (int64_t)chunks * (Kp / 16);
    const unsigned wgs = (unsigned)((waves + 3) / 4);
    const unsigned per = lgdhip::xcd_per(wgs), grid = lgdhip::xcd_groups(wgs);
    const int nt = (cout + 15) / 16;
#define LGD_WG_CASE(T)                                                                                       \
    case T:                                                                                                  \
        conv3d_wgrad_mfma<T><<<grid, 256, 0, s>>>((const u16 *)x, n, ldx, offx, cin, ksize, (const u16 *)dy, \
                                                  ldy, offy, cout, Kp, chunk_vox, P, per);                   \
        break;
    switch (nt) {
        LGD_WG_CASE(1)
        LGD_WG_CASE(2)
        LGD_WG_CASE(3)
        LGD_WG_CASE(4)
        LGD_WG_CASE(5)
        LGD_WG_CASE(6)
        LGD_WG_CASE(7)
        LGD_WG_CASE(8)
        LGD_WG_CASE(9)
        LGD_WG_CASE(10)
        LGD_WG_CASE(11)
        LGD_WG_CASE(12)
    }
#undef LGD_WG_CASE
    LGD_LAUNCHED("conv3d_wgrad_mfma");
    conv3d_wgrad_reduce<<<(unsigned)((cout * Kp + kWgradReduceOutputs - 1) / kWgradReduceOutputs),
                              dim3(kWgradReduceOutputs, kWgradReduceRows), 0, s>>>(P, chunks, cout, Kp, K, dw,
                                                                                    accumulate);
    LGD_LAUNCHED("conv3d_wgrad_reduce");
    if (db) {
        const int bchunks = (int)((nvox + kBiasChunk - 1) / kBiasChunk);
        float *Pb = P + (size_t)chunks * cout * Kp;
        bias_grad_partial<<<bchunks, 256, 0, s>>>((const u16 *)dy, nvox, ldy, offy, cout, Pb, bchunks);
        row_sum_final<<<blocks_for(cout), 256, 0, s>>>(Pb, bchunks, cout, db, accumulate);
        LGD_LAUNCHED("bias_grad");
    }
    return 0;
}

int lgdhip_conv3d_dgrad_weights(const void *w, int cin, int cout, int ksize, void *wt, lgdhip_stream_t stream) {
    LGD_REQUIRE(w && wt, "NULL pointer");
    LGD_REQUIRE(ksize == 1 || ksize == 3, "ksize must be 1 or 3");
    const int Kp = lgdhip_conv3d_kpad(cin, ksize);
    const int Kpt = lgdhip_conv3d_kpad(cout, ksize);
    conv3d_dgrad_weights<<<blocks_for((int64_t)cin * Kpt), 256, 0, lgdhip::as_stream(stream)>>>(
        (const u16 *)w, cin, cout, ksize, Kp, (u16 *)wt, Kpt);
    LGD_LAUNCHED("conv3d_dgrad_weights");
    return 0;
}

int lgdhip_relu_backward_bf16(void *g, int ldg, int offg, const void *y, int ldy, int offy, int c, int64_t nvox,
                              lgdhip_stream_t stream) {
    LGD_REQUIRE(g && y && c > 0 && nvox > 0, "bad arguments");
    relu_backward<<<blocks_for(nvox * c), 256, 0, lgdhip::as_stream(stream)>>>((u16 *)g, ldg, offg, (const u16 *)y,
                                                                               ldy, offy, c, nvox);
    LGD_LAUNCHED("relu_backward");
    return 0;
}

int lgdhip_avgpool2_bf16(const void *x, int n, int ldx, int offx, void *y, int ldy, int offy, int c,
                         lgdhip_stream_t stream) {
    LGD_REQUIRE(x && y && n % 2 == 0, "bad arguments");
    const int64_t m = n / 2;
    if (vec8(x, y, c, ldx, offx, ldy, offy) && n <= kMaxRowGrid) {
        avgpool2_v8<<<row_grid((int)m, c / 8), 256, 0, lgdhip::as_stream(stream)>>>(
            (const u16 *)x, n, ldx, offx, (u16 *)y, ldy, offy, c / 8);
        LGD_LAUNCHED("avgpool2_v8");
        return 0;
    }
    avgpool2<<<blocks_for(m * m * m * c), 256, 0, lgdhip::as_stream(stream)>>>((const u16 *)x, n, ldx, offx, (u16 *)y,
                                                                               ldy, offy, c);
    LGD_LAUNCHED("avgpool2");
    return 0;
}

int lgdhip_avgpool2_backward_bf16(const void *dy, int n, int ldy, int offy, void *dx, int ldx, int offx, int c,
                                  int accumulate, lgdhip_stream_t stream) {
    LGD_REQUIRE(dy && dx && n % 2 == 0, "bad arguments");
    if (vec8(dy, dx, c, ldy, offy, ldx, offx) && n <= kMaxRowGrid) {
        avgpool2_backward_v8<<<row_grid(n, c / 8), 256, 0, lgdhip::as_stream(stream)>>>(
            (const u16 *)dy, n, ldy, offy, (u16 *)dx, ldx, offx, c / 8, accumulate);
        LGD_LAUNCHED("avgpool2_backward_v8");
        return 0;
    }
    avgpool2_backward<<<blocks_for((int64_t)n * n * n * c), 256, 0, lgdhip::as_stream(stream)>>>(
        (const u16 *)dy, n, ldy, offy, (u16 *)dx, ldx, offx, c, accumulate);
    LGD_LAUNCHED("avgpool2_backward");
    return 0;
}

int lgdhip_upsample2_bf16(const void *x, int m, int ldx, int offx, void *y, int ldy, int offy, int c,
                          lgdhip_stream_t stream) {
    LGD_REQUIRE(x && y && m >= 1, "bad arguments");
    const int64_t n = 2 * (int64_t)m;
    if (vec8(x, y, c, ldx, offx, ldy, offy) && n <= kMaxRowGrid) {
        upsample2_v8<<<row_grid((int)n, c / 8), 256, 0, lgdhip::as_stream(stream)>>>(
            (const u16 *)x, m, ldx, offx, (u16 *)y, ldy, offy, c / 8);
        LGD_LAUNCHED("upsample2_v8");
        return 0;
    }
    upsample2<<<blocks_for(n * n * n * c), 256, 0, lgdhip::as_stream(stream)>>>((const u16 *)x, m, ldx, offx, (u16 *)y,
                                                                                ldy, offy, c);
    LGD_LAUNCHED("upsample2");
    return 0;
}

int lgdhip_upsample2_concat_bf16(const void *x, int m, int ldx, int offx, int cu, const void *skip, int lds,
                                 int offs, int cs, void *y, int ldy, int offy, lgdhip_stream_t stream) {
    LGD_REQUIRE(x && skip && y && m >= 1 && cu > 0 && cs > 0, "bad arguments");
    const int64_t n = 2 * (int64_t)m;
    if (vec8(x, y, cu, ldx, offx, ldy, offy) && vec8(skip, y, cs, lds, offs, ldy, offy) && n <= kMaxRowGrid) {
        upsample2_concat_v8<<<row_grid((int)n, (cu + cs) / 8), 256, 0, lgdhip::as_stream(stream)>>>(
            (const u16 *)x, m, ldx, offx, cu / 8, (const u16 *)skip, lds, offs, (u16 *)y, ldy, offy, (cu + cs) / 8);
        LGD_LAUNCHED("upsample2_concat_v8");
        return 0;
    }
    upsample2_concat<<<blocks_for(n * n * n * (cu + cs)), 256, 0, lgdhip::as_stream(stream)>>>(
        (const u16 *)x, m, ldx, offx, cu, (const u16 *)skip, lds, offs, cs, (u16 *)y, ldy, offy);
    LGD_LAUNCHED("upsample2_concat");
    return 0;
}

int lgdhip_upsample2_backward_bf16(const void *dy, int m, int ldy, int offy, void *dx, int ldx, int offx, int c,
                                   int accumulate, lgdhip_stream_t stream) {
    LGD_REQUIRE(dy && dx && m >= 1, "bad arguments");
    if (vec8(dy, dx, c, ldy, offy, ldx, offx) && 2 * (int64_t)m <= kMaxRowGrid) {
        upsample2_backward_v8<<<row_grid(m, c / 8), 256, 0, lgdhip::as_stream(stream)>>>(
            (const u16 *)dy, m, ldy, offy, (u16 *)dx, ldx, offx, c / 8, accumulate);
        LGD_LAUNCHED("upsample2_backward_v8");
        return 0;
    }
    upsample2_backward<<<blocks_for((int64_t)m * m * m * c), 256, 0, lgdhip::as_stream(stream)>>>(
        (const u16 *)dy, m, ldy, offy, (u16 *)dx, ldx, offx, c, accumulate);
    LGD_LAUNCHED("upsample2_backward");
    return 0;
}

int lgdhip_pack_bf16(const float *src, int64_t nvox, void *y, int ldy, int offy, lgdhip_stream_t stream) {
    LGD_REQUIRE(src && y && nvox > 0, "bad arguments");
    pack_bf16<<<blocks_for(nvox), 256, 0, lgdhip::as_stream(stream)>>>(src, nvox, (u16 *)y, ldy, offy);
    LGD_LAUNCHED("pack_bf16");
    return 0;
}

int64_t lgdhip_mse_partials(int64_t nvox) { return (nvox + 255) / 256; }

int lgdhip_mse_loss(const float *pred, const float *target, int64_t nvox, void *grad, int ldg, float *partial,
                    lgdhip_stream_t stream) {
    LGD_REQUIRE(pred && target && grad && partial && nvox > 0, "bad arguments");
    mse_loss<<<blocks_for(nvox), 256, 0, lgdhip::as_stream(stream)>>>(pred, target, nvox, (u16 *)grad, ldg, partial);
    LGD_LAUNCHED("mse_loss");
    return 0;
}

int lgdhip_adam_step(float *w, float *m, float *v, const float *g, int64_t count, int cols_used, int kp, float lr,
                     float beta1, float beta2, float eps, int t, void *w_bf16, lgdhip_stream_t stream) {
    LGD_REQUIRE(w && m && v && g && count > 0 && t >= 1, "bad arguments");
    const float c1 = 1.0f - powf(beta1, (float)t);
    const float c2 = 1.0f - powf(beta2, (float)t);
    adam_step<<<blocks_for(count), 256, 0, lgdhip::as_stream(stream)>>>(w, m, v, g, count, cols_used, kp, lr, beta1,
                                                                       beta2, eps, c1, c2, (u16 *)w_bf16);
    LGD_LAUNCHED("adam_step");
    return 0;
}

} // extern "C"
