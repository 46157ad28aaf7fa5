// Shared helpers of the lgdhip C-ABI implementation (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>

#include "../../include/lgdhip.h"

namespace lgdhip {

// Records a detailed message for lgdhip_last_error() and returns `code`.
int set_error(int code, const std::string &msg);

inline int hip_code(hipError_t e) { return e == hipSuccess ? 0 : LGDHIP_ERR_HIP - (int)e; }

int hip_fail(hipError_t e, const char *what);

#define LGD_HIP(call)                                              \
    do {                                                           \
        hipError_t lgd_e_ = (call);                                \
        if (lgd_e_ != hipSuccess) return ::lgdhip::hip_fail(lgd_e_, #call); \
    } while (0)

// Checks the launch status of the kernel just enqueued.
#define LGD_LAUNCHED(name)                                         \
    do {                                                           \
        hipError_t lgd_e_ = hipGetLastError();                     \
        if (lgd_e_ != hipSuccess) return ::lgdhip::hip_fail(lgd_e_, name); \
    } while (0)

#define LGD_REQUIRE(cond, msg)                                     \
    do {                                                           \
        if (!(cond)) return ::lgdhip::set_error(LGDHIP_ERR_INVALID_ARG, msg); \
    } while (0)

inline hipStream_t as_stream(lgdhip_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// compile-time int and loop: static_for<B, E>(f) calls f(ic<B>{}) .. f(ic<E-1>{})
template <int V>
using ic = std::integral_constant<int, V>;
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + 1, E>(f);
    }
}

// Validates the internal consistency of a descriptor (extents, offsets and
// pitches leave room for the halo).  Returns 0 or an error code.
// Views (LGDHIP_GRID_VIEW) are rejected unless allow_view.
int check_grid(const lgdhip_grid *g, int dim, int elem_bytes, int members, bool allow_view = false);

inline int64_t plane_elems(const lgdhip_grid &g) { return g.sx * g.sy * g.sz; }
// element strides between members and between z slices (LGDHIP_LAYOUT_*)
inline int64_t member_stride(const lgdhip_grid &g) {
    return g.layout == LGDHIP_LAYOUT_Z_SLABS ? g.sx * g.sy : g.sx * g.sy * g.sz;
}
inline int64_t z_stride(const lgdhip_grid &g) {
    return g.layout == LGDHIP_LAYOUT_Z_SLABS ? (int64_t)g.members * g.sx * g.sy : g.sx * g.sy;
}
// element index of logical (0, 0, 0) of member 0
inline int64_t origin_index(const lgdhip_grid &g) { return g.oz * z_stride(g) + g.oy * g.sx + g.ox; }

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// XCD-aware work distribution.  The dispatcher places work-group w on XCD
// w % 8 (8 XCDs, each with its own L2); tiles that share cache lines
// (overlap rings, rows straddled by shifted reads) should run on the same
// XCD at about the same time.  A launch of xcd_groups(T) work-groups gives
// XCD c the contiguous tile range [c*per, (c+1)*per), per = ceil(T/8), in
// order; ids mapped past T exit (before any barrier).
constexpr unsigned kXcds = 8;
// compute units of an MI355X (32 per XCD), for launch-size models
constexpr unsigned kCUs = 256;
inline unsigned xcd_per(unsigned tiles) { return (tiles + kXcds - 1) / kXcds; }
inline unsigned xcd_groups(unsigned tiles) { return kXcds * xcd_per(tiles); }
__device__ __forceinline__ unsigned xcd_tile(unsigned w, unsigned per, bool swizzle) {
    return swizzle ? (w % kXcds) * per + w / kXcds : w;
}
// Grouped order: groups of G consecutive tiles run on one XCD and
// consecutive groups on consecutive XCDs, so tiles sharing lines within a
// group share an L2 while the whole GPU still streams one contiguous front.
// Launch xcd_group_groups(T, G) work-groups; ids mapped past T exit.
inline unsigned xcd_group_groups(unsigned tiles, unsigned g) {
    const unsigned span = kXcds * g;
    return (tiles + span - 1) / span * span;
}
__device__ __forceinline__ unsigned xcd_group_tile(unsigned w, unsigned g) {
    const unsigned span = kXcds * g;
    return (w / span) * span + (w % kXcds) * g + (w / kXcds) % g;
}
// the swizzle choice of a tile's flags, else the family default
inline int xcd_swizzle(const lgdhip_tile *tile, bool family_default) {
    if (tile && (tile->flags & LGDHIP_TILE_NO_XCD_SWIZZLE)) return 0;
    if (tile && (tile->flags & LGDHIP_TILE_XCD_SWIZZLE)) return 1;
    return family_default ? 1 : 0;
}

} // namespace lgdhip
