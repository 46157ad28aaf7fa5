// Learned initial guess for the Poisson ensemble: data movement between the
// ensemble's padded pools, a coarse grid and the U-Net's dense tensors (no
// reference counterpart).
//
//  * coarse grid: problem n^3 (cells 0..n-1, Dirichlet halo at -1 and n) has
//    a coarse problem of m^3 cells with spacing H = (n+1)/(m+1), coarse cell
//    j at fine position X_j = (j+1) H - 1, so that the coarse halo j = -1,
//    j = m sits exactly on the fine halo planes -1, n (same faces);
//    restriction = hat-weighted average of f around X_j (weights
//    max(0, 1 - |i - X_j| / H) per axis), times H^2 (the coarse equation is
//    6U - sum U_nb = H^2 f_c); prolongation = trilinear interpolation of U at
//    fine position x, coarse coordinate (x+1)/H - 1, with coarse edge and
//    corner halo cells taken as the mean of the faces meeting there;
//  * network input: channel 0 = f / s, channel 1 = b / s (b: the face field,
//    each cell's sum of the Dirichlet values of the faces it touches),
//    channel 2 = base / s when a base field is given (dense n^3 FP32, x
//    fastest: e.g. the prolongated coarse solution), BF16, other channels 0;
//  * training target: (u - base) / s as FP32;  initial guess: u = s * pred +
//    base (FP64, halo untouched).
// Input and target can be read through one of the 48 symmetries of the cube
// (axis permutation x reflection), under which the 7-point problem -- and
// the coarse grid, which is symmetric in j -> m-1-j -- is invariant: a
// symmetric image of a solved problem is a solved problem.
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

__global__ void unet_input(const float *__restrict__ f, int n, double inv, Faces fc, int sym,
                           const float *__restrict__ base, u16 *__restrict__ X, int ldx) {
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
    if (base) p[c++] = f2bf((float)((double)base[((int64_t)sz * n + sy) * n + sx] * inv));
    for (; c < ldx; ++c) p[c] = 0;
}

__global__ void unet_target(const double *__restrict__ u, int n, double inv, int sym, const float *__restrict__ base,
                            float *__restrict__ T) {
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    int sx, sy, sz;
    sym_source(sym, n, x, y, z, sx, sy, sz);
    const int64_t np = n + 2;
    double w = u[((int64_t)(sz + 1) * np + (sy + 1)) * np + (sx + 1)];
    if (base) w -= (double)base[((int64_t)sz * n + sy) * n + sx];
    T[v] = (float)(w * inv);
}

__global__ void unet_guess(const float *__restrict__ pred, int n, double scale, const float *__restrict__ base,
                           double *__restrict__ u) {
#pragma clang fp contract(off)
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
    const int64_t np = n + 2;
    double w = pred ? (double)pred[v] * scale : 0.0;
    if (base) w += (double)base[v];
    u[((int64_t)(z + 1) * np + (y + 1)) * np + (x + 1)] = w;
}

// f_c(J) = H^2 * sum_i w(i,J) f(i) / sum_i w(i,J), w = product over axes of
// max(0, 1 - |i - X_J| / H), X_J = (J+1) H - 1, i over the fine interior
__global__ void coarse_restrict(const float *__restrict__ f, int n, int m, double H, float *__restrict__ fc) {
    const int64_t ncell = (int64_t)m * m * m;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= ncell) return;
    const int jx = (int)(v % m), jy = (int)((v / m) % m), jz = (int)(v / ((int64_t)m * m));
    const double X[3] = {(jx + 1) * H - 1.0, (jy + 1) * H - 1.0, (jz + 1) * H - 1.0};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = max(0, (int)ceil(X[a] - H));
        hi[a] = min(n - 1, (int)floor(X[a] + H));
    }
    const int64_t np = n + 2;
    double acc = 0.0, wsum = 0.0;
    for (int z = lo[2]; z <= hi[2]; ++z) {
        const double wz = fmax(0.0, 1.0 - fabs(z - X[2]) / H);
        for (int y = lo[1]; y <= hi[1]; ++y) {
            const double wzy = wz * fmax(0.0, 1.0 - fabs(y - X[1]) / H);
            const float *row = f + ((int64_t)(z + 1) * np + (y + 1)) * np + 1;
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const double w = wzy * fmax(0.0, 1.0 - fabs(x - X[0]) / H);
                acc += w * (double)row[x];
                wsum += w;
            }
        }
    }
    const int64_t mp = m + 2;
    fc[((int64_t)(jz + 1) * mp + (jy + 1)) * mp + (jx + 1)] = (float)(wsum > 0.0 ? H * H * acc / wsum : 0.0);
}

// coarse value at padded index (a, b, c) in [0, m+1]^3: interior and face
// cells as stored, edge / corner halo cells the mean of the faces meeting
// there
__device__ __forceinline__ double coarse_at(const double *__restrict__ U, int m, const Faces &fc, int a, int b,
                                            int c) {
    const int hx = a == 0 ? 0 : (a == m + 1 ? 1 : -1);
    const int hy = b == 0 ? 2 : (b == m + 1 ? 3 : -1);
    const int hz = c == 0 ? 4 : (c == m + 1 ? 5 : -1);
    const int nh = (hx >= 0) + (hy >= 0) + (hz >= 0);
    const int64_t mp = m + 2;
    if (nh <= 1) return U[((int64_t)c * mp + b) * mp + a];
    double s = 0.0;
    if (hx >= 0) s += fc.v[hx];
    if (hy >= 0) s += fc.v[hy];
    if (hz >= 0) s += fc.v[hz];
    return s / nh;
}

__global__ void coarse_prolong(const double *__restrict__ U, int m, double H, Faces fc, int n,
                               float *__restrict__ out) {
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int c3[3] = {(int)(v % n), (int)((v / n) % n), (int)(v / ((int64_t)n * n))};
    // padded coarse coordinate (x+1)/H in (0, m+1): cells k, k+1 of the
    // padded block (0 and m+1 are the halo)
    int i0[3];
    double t[3];
    for (int a = 0; a < 3; ++a) {
        const double xi = (c3[a] + 1) / H;
        int k = (int)floor(xi);
        k = min(max(k, 0), m);
        i0[a] = k;
        t[a] = xi - k;
    }
    double s = 0.0;
    for (int d = 0; d < 8; ++d) {
        const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
        const double w = (dx ? t[0] : 1.0 - t[0]) * (dy ? t[1] : 1.0 - t[1]) * (dz ? t[2] : 1.0 - t[2]);
        s += w * coarse_at(U, m, fc, i0[0] + dx, i0[1] + dy, i0[2] + dz);
    }
    out[v] = (float)s;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

int check_problem(const lgdhip_poisson_problem *prob) {
    if (!prob || prob->n < 1 || prob->offset < 0)
        return lgdhip::set_error(LGDHIP_ERR_INVALID_ARG, "bad Poisson problem descriptor");
    return 0;
}

Faces faces_of(const lgdhip_poisson_problem *prob) {
    Faces fc;
    for (int j = 0; j < 6; ++j) fc.v[j] = prob->face[j];
    return fc;
}

} // namespace

extern "C" {

int lgdhip_poisson_unet_input(const float *f, const lgdhip_poisson_problem *prob, double inv_scale, int symmetry,
                              const float *base, void *x, int ldx, lgdhip_stream_t stream) {
    LGD_REQUIRE(f && x, "NULL pointer");
    LGD_REQUIRE(ldx >= (base ? 3 : 2), "ldx must be >= 2 (3 with a base field)");
    LGD_REQUIRE(symmetry >= 0 && symmetry < 48, "symmetry must be in [0, 48)");
    int rc = check_problem(prob);
    if (rc) return rc;
    const int n = prob->n;
    unet_input<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(
        f + prob->offset, n, inv_scale, faces_of(prob), symmetry, base, (u16 *)x, ldx);
    LGD_LAUNCHED("unet_input");
    return 0;
}

int lgdhip_poisson_unet_target(const double *u, const lgdhip_poisson_problem *prob, double inv_scale, int symmetry,
                               const float *base, float *t, lgdhip_stream_t stream) {
    LGD_REQUIRE(u && t, "NULL pointer");
    LGD_REQUIRE(symmetry >= 0 && symmetry < 48, "symmetry must be in [0, 48)");
    int rc = check_problem(prob);
    if (rc) return rc;
    const int n = prob->n;
    unet_target<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(u + prob->offset, n, inv_scale,
                                                                                      symmetry, base, t);
    LGD_LAUNCHED("unet_target");
    return 0;
}

int lgdhip_poisson_unet_guess(const float *pred, const lgdhip_poisson_problem *prob, double scale, const float *base,
                              double *u, lgdhip_stream_t stream) {
    LGD_REQUIRE(u && (pred || base), "NULL pointer");
    int rc = check_problem(prob);
    if (rc) return rc;
    const int n = prob->n;
    unet_guess<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(pred, n, scale, base,
                                                                                     u + prob->offset);
    LGD_LAUNCHED("unet_guess");
    return 0;
}

double lgdhip_poisson_coarse_spacing(int n, int m) { return (n >= 1 && m >= 1) ? (double)(n + 1) / (m + 1) : 0.0; }

int lgdhip_poisson_coarse_restrict(const float *f, const lgdhip_poisson_problem *prob, float *fc,
                                   const lgdhip_poisson_problem *coarse, lgdhip_stream_t stream) {
    LGD_REQUIRE(f && fc, "NULL pointer");
    int rc = check_problem(prob);
    if (rc) return rc;
    rc = check_problem(coarse);
    if (rc) return rc;
    const int n = prob->n, m = coarse->n;
    LGD_REQUIRE(m <= n, "the coarse grid must not be finer than the problem");
    coarse_restrict<<<blocks_for((int64_t)m * m * m), 256, 0, lgdhip::as_stream(stream)>>>(
        f + prob->offset, n, m, lgdhip_poisson_coarse_spacing(n, m), fc + coarse->offset);
    LGD_LAUNCHED("coarse_restrict");
    return 0;
}

int lgdhip_poisson_coarse_prolong(const double *uc, const lgdhip_poisson_problem *coarse, int n, float *out,
                                  lgdhip_stream_t stream) {
    LGD_REQUIRE(uc && out, "NULL pointer");
    int rc = check_problem(coarse);
    if (rc) return rc;
    const int m = coarse->n;
    LGD_REQUIRE(n >= m, "the fine grid must not be coarser than the coarse problem");
    coarse_prolong<<<blocks_for((int64_t)n * n * n), 256, 0, lgdhip::as_stream(stream)>>>(
        uc + coarse->offset, m, lgdhip_poisson_coarse_spacing(n, m), faces_of(coarse), n, out);
    LGD_LAUNCHED("coarse_prolong");
    return 0;
}

} // extern "C"
