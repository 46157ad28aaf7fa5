// Data movement between the Poisson ensemble's padded pools and the U-Net's
// dense channels-last tensors (no reference counterpart):
//  * network input: channel 0 = f / s, channel 1 = b / s, where b is the
//    face field (each cell's sum of the Dirichlet values of the faces it
//    touches: the boundary part of the zero-guess residual), optionally
//    channel 2 = h / s with h = sum_j face_j H_j the exact solution of the
//    boundary part (see below), BF16, remaining channels zero;
//  * the harmonic basis H: the n^3 solution (x fastest) of the problem with
//    f = 0, value 1 on the face x = -1 and 0 on the others; by the cube's
//    symmetry the basis field of face j is H read at
//      j=0 (x=-1): (x, y, z)      j=1 (x=n): (n-1-x, y, z)
//      j=2 (y=-1): (y, x, z)      j=3 (y=n): (n-1-y, x, z)
//      j=4 (z=-1): (z, y, x)      j=5 (z=n): (n-1-z, y, x)
//    and by linearity sum_j face_j H_j solves f = 0 with the problem's faces;
//  * training target: interior of u / s as FP32, or of (u - h) / s (the
//    source response alone) when H is given;
//  * initial guess: prediction * s (+ h when H is given) written into the
//    interior of u (FP64).
// Input and target can be read through one of the 48 symmetries of the cube
// (axis permutation x reflection), under which the 7-point problem is
// invariant: a symmetric image of a solved problem is a solved problem.
#include "lgdhip_internal.h"

namespace {

typedef unsigned short u16;

__device__ __forceinline__ u16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}

struct Faces {
    double v[6];
};

// symmetry t = perm + 6 * flips: source coordinate s[perm_i] = o[i], then
// s[a] = n-1-s[a] for each axis a whose bit is set in flips
__device__ __forceinline__ void sym_source(int t, int n, int x, int y, int z, int &sx, int &sy, int &sz) {
    const int p = t % 6, fl = t / 6;
    // perms of (0,1,2): 012 021 102 120 201 210; o = (x, y, z)
    const int p0 = (p < 2) ? 0 : (p < 4 ? 1 : 2);
    const int p1 = (p == 0 || p == 5) ? 1 : ((p == 1 || p == 3) ? 2 : 0);
    const int p2 = 3 - p0 - p1;
    int s[3];
    s[p0] = x;
    s[p1] = y;
    s[p2] = z;
    if (fl & 1) s[0] = n - 1 - s[0];
    if (fl & 2) s[1] = n - 1 - s[1];
    if (fl & 4) s[2] = n - 1 - s[2];
    sx = s[0];
    sy = s[1];
    sz = s[2];
}

// h(x,y,z) = sum_j face_j H_j(x,y,z), summed j = 0..5 in FP64 (separate
// multiplies and adds: no FMA contraction, so the value is reproducible)
__device__ __forceinline__ double face_solution(const float *__restrict__ H, int n, const Faces &fc, int x, int y,
                                                int z) {
#pragma clang fp contract(off)
    auto at = [&](int a, int b, int c) { return (double)H[((int64_t)c * n + b) * n + a]; };
    double h = fc.v[0] * at(x, y, z);
    h += fc.v[1] * at(n - 1 - x, y, z);
    h += fc.v[2] * at(y, x, z);
    h += fc.v[3] * at(n - 1 - y, x, z);
    h += fc.v[4] * at(z, y, x);
    h += fc.v[5] * at(n - 1 - z, y, x);
    return h;
}

__global__ void unet_input(const float *__restrict__ f, int n, double inv, Faces fc, int sym,
                           const float *__restrict__ H, u16 *__restrict__ X, int ldx) {
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    int sx, sy, sz;
    sym_source(sym, n, x, y, z, sx, sy, sz);
    const int64_t np = n + 2;
    const float fv = f[((int64_t)(sz + 1) * np + (sy + 1)) * np + (sx + 1)];
    double b = 0.0;
    if (sx == 0) b += fc.v[0];
    if (sx == n - 1) b += fc.v[1];
    if (sy == 0) b += fc.v[2];
    if (sy == n - 1) b += fc.v[3];
    if (sz == 0) b += fc.v[4];
    if (sz == n - 1) b += fc.v[5];
    u16 *p = X + v * ldx;
    p[0] = f2bf((float)((double)fv * inv));
    p[1] = f2bf((float)(b * inv));
    int c = 2;
    if (H) p[c++] = f2bf((float)(face_solution(H, n, fc, sx, sy, sz) * inv));
    for (; c < ldx; ++c) p[c] = 0;
}

__global__ void face_guess(const float *__restrict__ H, int n, Faces fc, double *__restrict__ u) {
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    const int64_t np = n + 2;
    u[((int64_t)(z + 1) * np + (y + 1)) * np + (x + 1)] = face_solution(H, n, fc, x, y, z);
}

__global__ void unet_target(const double *__restrict__ u, int n, double inv, int sym, const float *__restrict__ H,
                            Faces fc, float *__restrict__ T) {
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    int sx, sy, sz;
    sym_source(sym, n, x, y, z, sx, sy, sz);
    const int64_t np = n + 2;
    double w = u[((int64_t)(sz + 1) * np + (sy + 1)) * np + (sx + 1)];
    if (H) w -= face_solution(H, n, fc, sx, sy, sz);
    T[v] = (float)(w * inv);
}

__global__ void unet_guess(const float *__restrict__ pred, int n, double scale, const float *__restrict__ H, Faces fc,
                           double *__restrict__ u) {
#pragma clang fp contract(off)
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    const int64_t np = n + 2;
    double w = (double)pred[v] * scale;
    if (H) w += face_solution(H, n, fc, x, y, z);
    u[((int64_t)(z + 1) * np + (y + 1)) * np + (x + 1)] = w;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

int check_problem(const lgdhip_poisson_problem *prob) {
    if (!prob || prob->n < 1 || prob->offset < 0)
        return lgdhip::set_error(LGDHIP_ERR_INVALID_ARG, "bad Poisson problem descriptor");
    return 0;
}

} // namespace

extern "C" {

int lgdhip_poisson_unet_input(const float *f, const lgdhip_poisson_problem *prob, double inv_scale, int symmetry,
                              const float *harmonic, void *x, int ldx, lgdhip_stream_t stream) {
    LGD_REQUIRE(f && x, "NULL pointer");
    LGD_REQUIRE(ldx >= (harmonic ? 3 : 2), "ldx must be >= 2 (3 with the harmonic channel)");
    LGD_REQUIRE(symmetry >= 0 && symmetry < 48, "symmetry must be in [0, 48)");
    int rc = check_problem(prob);
    if (rc) return rc;
    Faces fc;
    for (int j = 0; j < 6; ++j) fc.v[j] = prob->face[j];
    const int n = prob->n;
    unet_input<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(f + prob->offset, n, inv_scale,
                                                                                     fc, symmetry, harmonic,
                                                                                     (u16 *)x, ldx);
    LGD_LAUNCHED("unet_input");
    return 0;
}

int lgdhip_poisson_unet_target(const double *u, const lgdhip_poisson_problem *prob, double inv_scale, int symmetry,
                               const float *harmonic, float *t, lgdhip_stream_t stream) {
    LGD_REQUIRE(u && t, "NULL pointer");
    LGD_REQUIRE(symmetry >= 0 && symmetry < 48, "symmetry must be in [0, 48)");
    int rc = check_problem(prob);
    if (rc) return rc;
    Faces fc;
    for (int j = 0; j < 6; ++j) fc.v[j] = prob->face[j];
    const int n = prob->n;
    unet_target<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(u + prob->offset, n, inv_scale,
                                                                                      symmetry, harmonic, fc, t);
    LGD_LAUNCHED("unet_target");
    return 0;
}

int lgdhip_poisson_unet_guess(const float *pred, const lgdhip_poisson_problem *prob, double scale,
                              const float *harmonic, double *u, lgdhip_stream_t stream) {
    LGD_REQUIRE(pred && u, "NULL pointer");
    int rc = check_problem(prob);
    if (rc) return rc;
    Faces fc;
    for (int j = 0; j < 6; ++j) fc.v[j] = prob->face[j];
    const int n = prob->n;
    unet_guess<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(pred, n, scale, harmonic, fc,
                                                                                     u + prob->offset);
    LGD_LAUNCHED("unet_guess");
    return 0;
}

int lgdhip_poisson_face_guess(const float *harmonic, const lgdhip_poisson_problem *prob, double *u,
                              lgdhip_stream_t stream) {
    LGD_REQUIRE(harmonic && u, "NULL pointer");
    int rc = check_problem(prob);
    if (rc) return rc;
    Faces fc;
    for (int j = 0; j < 6; ++j) fc.v[j] = prob->face[j];
    const int n = prob->n;
    face_guess<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(harmonic, n, fc,
                                                                                     u + prob->offset);
    LGD_LAUNCHED("face_guess");
    return 0;
}

} // extern "C"
