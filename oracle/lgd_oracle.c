/*
 * lgd_oracle.c -- CPU restatement of the reference's serial sweep semantics
 * for the kernel families of liblgdhip.  TEST INFRASTRUCTURE ONLY: it is the
 * checker that tests/, the smoke entry point and the CPU-baseline leg of the
 * benchmark compare the HIP kernels against; no product code path calls it.
 *
 * What is restated (paths relative to the libgeodecomp tree):
 *  - the SerialSimulator step loop with double buffering and swap
 *    (src/libgeodecomp/parallelization/serialsimulator.h:76-93, :132-139);
 *  - Cube topology: out-of-bounds neighbours read the grid's constant edge
 *    cell (src/libgeodecomp/geometry/topologies.h:186-199, storage/grid.h
 *    edge cell);
 *  - the user cells' update rules:
 *      JacobiCellFixedHood   src/testbed/performancetests/main.cpp:1194-1205
 *      ConwayCell            src/examples/gameoflife/main.cpp:36-59
 *      D3Q19 BGK cell        src/examples/latticeboltzmann/main.cpp:92-229
 *    (LBM with the cavity boundary described in lgd_oracle.h).
 *
 * Parity pinning of every function is listed in lgd_oracle.h.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "lgd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Jacobi3D FP64                                                       */
/* ------------------------------------------------------------------ */
/* periodic neighbour index along one axis: the opposite face's cell when
 * the axis wraps (Topologies::Torus normalize, geometry/topologies.h:225-247) */
static long wrap_prev(long c, long n) { return c > 0 ? c - 1 : n - 1; }
static long wrap_next(long c, long n) { return c < n - 1 ? c + 1 : 0; }

static void jacobi_sweep(const double *src, double *dst, long nx, long ny, long nz, double edge, int wrap,
                         int threads) {
    const long sxy = nx * ny;
    const int wx = wrap & 1, wy = wrap & 2, wz = wrap & 4;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (long z = 0; z < nz; ++z) {
        for (long y = 0; y < ny; ++y) {
            for (long x = 0; x < nx; ++x) {
                const long i = z * sxy + y * nx + x;
                const long row = z * sxy + y * nx, col = y * nx + x;
                /* hood[FixedCoord<...>] in the order of main.cpp:1196-1202;
                 * a neighbour past an unwrapped face is the edge cell */
                const double zm = z > 0 ? src[i - sxy] : wz ? src[wrap_prev(z, nz) * sxy + col] : edge;
                const double ym = y > 0 ? src[i - nx] : wy ? src[z * sxy + wrap_prev(y, ny) * nx + x] : edge;
                const double xm = x > 0 ? src[i - 1] : wx ? src[row + wrap_prev(x, nx)] : edge;
                const double c = src[i];
                const double xp = x < nx - 1 ? src[i + 1] : wx ? src[row + wrap_next(x, nx)] : edge;
                const double yp = y < ny - 1 ? src[i + nx] : wy ? src[z * sxy + wrap_next(y, ny) * nx + x] : edge;
                const double zp = z < nz - 1 ? src[i + sxy] : wz ? src[wrap_next(z, nz) * sxy + col] : edge;
                dst[i] = (zm + ym + xm + c + xp + yp + zp) * (1.0 / 7.0);
            }
        }
    }
}

int lgdo_jacobi3d_f64_wrap(double *grid, long nx, long ny, long nz, double edge, int steps, int threads, int wrap) {
    if (nx < 1 || ny < 1 || nz < 1 || steps < 0 || wrap < 0 || wrap > 7) return -1;
    const size_t n = (size_t)nx * ny * nz;
    double *tmp = (double *)malloc(n * sizeof(double));
    if (!tmp) return -2;
    double *cur = grid, *nxt = tmp;
    for (int t = 0; t < steps; ++t) {
        jacobi_sweep(cur, nxt, nx, ny, nz, edge, wrap, threads);
        double *s = cur;
        cur = nxt;
        nxt = s;
    }
    if (cur != grid) memcpy(grid, cur, n * sizeof(double));
    free(tmp);
    return 0;
}

int lgdo_jacobi3d_f64(double *grid, long nx, long ny, long nz, double edge, int steps, int threads) {
    return lgdo_jacobi3d_f64_wrap(grid, nx, ny, nz, edge, steps, threads, 0);
}

/* ------------------------------------------------------------------ */
/* Game of Life                                                        */
/* ------------------------------------------------------------------ */
static void life_sweep(const uint8_t *src, uint8_t *dst, long nx, long ny, int edge, int wrap, int threads) {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (long y = 0; y < ny; ++y) {
        for (long x = 0; x < nx; ++x) {
            /* countLivingNeighbors: all nine cells minus the centre; a
             * coordinate past a wrapped face is taken modulo the extent
             * (Topologies::Torus normalize, geometry/topologies.h:225-247) */
            int n = 0;
            for (int dy = -1; dy < 2; ++dy) {
                for (int dx = -1; dx < 2; ++dx) {
                    long xx = x + dx, yy = y + dy;
                    if (wrap & 1) xx = (xx + nx) % nx;
                    if (wrap & 2) yy = (yy + ny) % ny;
                    const int alive = (xx < 0 || yy < 0 || xx >= nx || yy >= ny) ? edge : (src[yy * nx + xx] != 0);
                    n += alive;
                }
            }
            const int self = src[y * nx + x] != 0;
            n -= self;
            int next;
            if (self) {
                next = (2 <= n) && (n <= 3);
            } else {
                next = (n == 3);
            }
            dst[y * nx + x] = (uint8_t)next;
        }
    }
}

int lgdo_life2d_u8_wrap(uint8_t *grid, long nx, long ny, int edge, int steps, int threads, int wrap) {
    if (nx < 1 || ny < 1 || steps < 0 || (edge != 0 && edge != 1) || wrap < 0 || wrap > 3) return -1;
    const size_t n = (size_t)nx * ny;
    uint8_t *tmp = (uint8_t *)malloc(n);
    if (!tmp) return -2;
    uint8_t *cur = grid, *nxt = tmp;
    for (int t = 0; t < steps; ++t) {
        life_sweep(cur, nxt, nx, ny, edge, wrap, threads);
        uint8_t *s = cur;
        cur = nxt;
        nxt = s;
    }
    if (cur != grid) memcpy(grid, cur, n);
    free(tmp);
    return 0;
}

int lgdo_life2d_u8(uint8_t *grid, long nx, long ny, int edge, int steps, int threads) {
    return lgdo_life2d_u8_wrap(grid, nx, ny, edge, steps, threads, 0);
}

/* ------------------------------------------------------------------ */
/* D3Q19 lid-driven cavity                                             */
/* ------------------------------------------------------------------ */
/* C N E W S T B NW SW NE SE TW BW TE BE TN BN TS BS
 * (latticeboltzmann/main.cpp:36-55 and the pull offsets of :106-174) */
static const int EX[19] = {0, 0, 1, -1, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0};
static const int EY[19] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
static const int OPP[19] = {0, 4, 3, 2, 1, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

static double weight_of(int i) { return i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }

int lgdo_lbm_direction(int i, int *ex, int *ey, int *ez, int *opp) {
    if (i < 0 || i >= 19) return -1;
    *ex = EX[i];
    *ey = EY[i];
    *ez = EZ[i];
    *opp = OPP[i];
    return 0;
}

static int link_blocked(int i, unsigned fl) {
    return (EX[i] == 1 && (fl & LGDO_WALL_XLO)) || (EX[i] == -1 && (fl & LGDO_WALL_XHI)) ||
           (EY[i] == 1 && (fl & LGDO_WALL_YLO)) || (EY[i] == -1 && (fl & LGDO_WALL_YHI)) ||
           (EZ[i] == 1 && (fl & LGDO_WALL_ZLO)) || (EZ[i] == -1 && (fl & LGDO_WALL_ZHI));
}

void lgdo_lbm_cavity_flags(uint8_t *flags, long nx, long ny, long nz) {
    for (long z = 0; z < nz; ++z)
        for (long y = 0; y < ny; ++y)
            for (long x = 0; x < nx; ++x) {
                unsigned fl = 0;
                if (x == 0) fl |= LGDO_WALL_XLO;
                if (x == nx - 1) fl |= LGDO_WALL_XHI;
                if (y == 0) fl |= LGDO_WALL_YLO;
                if (y == ny - 1) fl |= LGDO_WALL_YHI | LGDO_LID_YHI;
                if (z == 0) fl |= LGDO_WALL_ZLO;
                if (z == nz - 1) fl |= LGDO_WALL_ZHI;
                flags[(z * ny + y) * nx + x] = (uint8_t)fl;
            }
}

/* One cell of the fused pull stream + collide, written once for double and
 * once for float (same operation order). */
#define LGDO_LBM_CELL(T)                                                                              \
    static void lbm_cell_##T(const T *src, T *dst, const uint8_t *flags, long nx, long ny, long nz,     \
                             long x, long y, long z, T omega, T lux, T luy, T luz) {                   \
        const long plane = nx * ny * nz;                                                               \
        const long idx = (z * ny + y) * nx + x;                                                        \
        const unsigned fl = flags[idx];                                                                \
        T f[19];                                                                                       \
        for (int i = 0; i < 19; ++i) {                                                                 \
            if (link_blocked(i, fl)) {                                                                 \
                T v = src[OPP[i] * plane + idx];                                                       \
                if (EY[i] == -1 && (fl & LGDO_WALL_YHI) && (fl & LGDO_LID_YHI))                        \
                    v += (T)6 * (T)weight_of(i) * ((T)EX[i] * lux + (T)EY[i] * luy + (T)EZ[i] * luz);  \
                f[i] = v;                                                                              \
            } else {                                                                                   \
                const long sx_ = x - EX[i], sy_ = y - EY[i], sz_ = z - EZ[i];                          \
                f[i] = src[i * plane + (sz_ * ny + sy_) * nx + sx_];                                   \
            }                                                                                          \
        }                                                                                              \
        T rho = 0, jx = 0, jy = 0, jz = 0;                                                             \
        for (int i = 0; i < 19; ++i) {                                                                 \
            rho += f[i];                                                                               \
            if (EX[i]) jx += (T)EX[i] * f[i];                                                          \
            if (EY[i]) jy += (T)EY[i] * f[i];                                                          \
            if (EZ[i]) jz += (T)EZ[i] * f[i];                                                          \
        }                                                                                              \
        const T dterm = rho * ((T)1 / (T)3) - (T)0.5 * (jx * jx + jy * jy + jz * jz);                  \
        const T keep = (T)1 - omega;                                                                   \
        for (int i = 0; i < 19; ++i) {                                                                 \
            const T eu = (T)EX[i] * jx + (T)EY[i] * jy + (T)EZ[i] * jz;                                \
            const T w3 = (T)3 * (T)weight_of(i) * omega;                                               \
            dst[i * plane + idx] = keep * f[i] + w3 * (dterm + eu + (T)1.5 * eu * eu);                 \
        }                                                                                              \
    }

LGDO_LBM_CELL(double)
LGDO_LBM_CELL(float)

#define LGDO_LBM_RUN(T)                                                                                 \
    int lgdo_lbm_d3q19_##T(T *f, const uint8_t *flags, long nx, long ny, long nz, T omega, T lux, T luy, \
                           T luz, int steps, int threads) {                                              \
        if (nx < 1 || ny < 1 || nz < 1 || steps < 0) return -1;                                          \
        const size_t n = (size_t)19 * nx * ny * nz;                                                      \
        T *tmp = (T *)malloc(n * sizeof(T));                                                             \
        if (!tmp) return -2;                                                                             \
        T *cur = f, *nxt = tmp;                                                                          \
        for (int t = 0; t < steps; ++t) {                                                                \
            _Pragma("omp parallel for schedule(static) num_threads(threads) if (threads > 1)")            \
            for (long z = 0; z < nz; ++z)                                                                \
                for (long y = 0; y < ny; ++y)                                                            \
                    for (long x = 0; x < nx; ++x)                                                        \
                        lbm_cell_##T(cur, nxt, flags, nx, ny, nz, x, y, z, omega, lux, luy, luz);        \
            T *s = cur;                                                                                  \
            cur = nxt;                                                                                   \
            nxt = s;                                                                                     \
        }                                                                                                \
        if (cur != f) memcpy(f, cur, n * sizeof(T));                                                     \
        free(tmp);                                                                                       \
        return 0;                                                                                        \
    }

LGDO_LBM_RUN(double)
LGDO_LBM_RUN(float)

void lgdo_lbm_equilibrium_f64(double *f, const double *rho, const double *jx, const double *jy, const double *jz,
                              long ncells) {
    for (long c = 0; c < ncells; ++c) {
        const double r = rho[c], ux = jx[c], uy = jy[c], uz = jz[c];
        const double dterm = r * (1.0 / 3.0) - 0.5 * (ux * ux + uy * uy + uz * uz);
        for (int i = 0; i < 19; ++i) {
            const double eu = EX[i] * ux + EY[i] * uy + EZ[i] * uz;
            f[i * ncells + c] = 3.0 * weight_of(i) * (dterm + eu + 1.5 * eu * eu);
        }
    }
}

void lgdo_lbm_moments_f64(const double *f, long ncells, double *rho, double *jx, double *jy, double *jz) {
    for (long c = 0; c < ncells; ++c) {
        double r = 0, x = 0, y = 0, z = 0;
        for (int i = 0; i < 19; ++i) {
            const double v = f[i * ncells + c];
            r += v;
            if (EX[i]) x += EX[i] * v;
            if (EY[i]) y += EY[i] * v;
            if (EZ[i]) z += EZ[i] * v;
        }
        rho[c] = r;
        jx[c] = x;
        jy[c] = y;
        jz[c] = z;
    }
}

int lgdo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Poisson problems (padded (n+2)^3 blocks, x fastest)                 */
/* ------------------------------------------------------------------ */
void lgdo_poisson_set_faces(double *u, long n, const double face[6]) {
    const long s = n + 2;
    for (long b = 1; b <= n; ++b)
        for (long a = 1; a <= n; ++a) {
            u[(b * s + a) * s + 0] = face[0];
            u[(b * s + a) * s + (n + 1)] = face[1];
            u[(b * s + 0) * s + a] = face[2];
            u[(b * s + (n + 1)) * s + a] = face[3];
            u[(0 * s + b) * s + a] = face[4];
            u[((n + 1) * s + b) * s + a] = face[5];
        }
}

/* one weighted (damped) FP32 Jacobi sweep e' = e + omega (J(e) - e): the
 * plain sweep's sum, then e' = fma(omega, fma(sum, 1/6, -e), e) with single
 * roundings (C fmaf, correctly rounded), the device kernels' form */
void lgdo_poisson_sweep_damped_f32(const float *e, float *e2, const float *r, long n, float omega) {
    const long s = n + 2, sxy = s * s;
    for (long z = 1; z <= n; ++z)
        for (long y = 1; y <= n; ++y)
            for (long x = 1; x <= n; ++x) {
                const long i = (z * s + y) * s + x;
                float v = e[i - sxy];
                v += e[i - s];
                v += e[i - 1];
                v += e[i + 1];
                v += e[i + s];
                v += e[i + sxy];
                v += r[i];
                e2[i] = fmaf(omega, fmaf(v, 1.0f / 6.0f, -e[i]), e[i]);
            }
}

/* one FP32 Jacobi sweep on A e = r, in the summation order of the kernel */
void lgdo_poisson_sweep_f32(const float *e, float *e2, const float *r, long n) {
    const long s = n + 2, sxy = s * s;
    for (long z = 1; z <= n; ++z)
        for (long y = 1; y <= n; ++y)
            for (long x = 1; x <= n; ++x) {
                const long i = (z * s + y) * s + x;
                float v = e[i - sxy];
                v += e[i - s];
                v += e[i - 1];
                v += e[i + 1];
                v += e[i + s];
                v += e[i + sxy];
                v += r[i];
                e2[i] = v * (1.0f / 6.0f);
            }
}

double lgdo_poisson_residual_norm(const double *u, const float *f, long n) {
    const long s = n + 2, sxy = s * s;
    double acc = 0.0;
    for (long z = 1; z <= n; ++z)
        for (long y = 1; y <= n; ++y)
            for (long x = 1; x <= n; ++x) {
                const long i = (z * s + y) * s + x;
                const double nb = u[i - sxy] + u[i - s] + u[i - 1] + u[i + 1] + u[i + s] + u[i + sxy];
                const double r = (double)f[i] + nb - 6.0 * u[i];
                acc += r * r;
            }
    return sqrt(acc);
}

/* plain FP64 Jacobi u <- (sum_nb u + f)/6 from the given u until
 * ||r|| <= rel_tol * ||r(initial u)|| (or reference if > 0), checking every
 * `check` sweeps; returns the sweep count */
long lgdo_poisson_jacobi_f64(double *u, const float *f, long n, double rel_tol, double reference, long max_sweeps,
                             long check, double *final_norm) {
    const long s = n + 2, sxy = s * s, total = s * s * s;
    double *v = (double *)malloc((size_t)total * sizeof(double));
    if (!v) return -1;
    memcpy(v, u, (size_t)total * sizeof(double));
    const double ref = reference > 0 ? reference : lgdo_poisson_residual_norm(u, f, n);
    double *cur = u, *nxt = v;
    long sweeps = 0;
    double rn = lgdo_poisson_residual_norm(cur, f, n);
    while (rn > rel_tol * ref && sweeps < max_sweeps) {
        for (long k = 0; k < check; ++k) {
            for (long z = 1; z <= n; ++z)
                for (long y = 1; y <= n; ++y)
                    for (long x = 1; x <= n; ++x) {
                        const long i = (z * s + y) * s + x;
                        const double nb = cur[i - sxy] + cur[i - s] + cur[i - 1] + cur[i + 1] + cur[i + s] + cur[i + sxy];
                        nxt[i] = (nb + (double)f[i]) * (1.0 / 6.0);
                    }
            double *t = cur;
            cur = nxt;
            nxt = t;
            ++sweeps;
        }
        rn = lgdo_poisson_residual_norm(cur, f, n);
    }
    if (cur != u) memcpy(u, cur, (size_t)total * sizeof(double));
    free(v);
    if (final_norm) *final_norm = rn;
    return sweeps;
}

/* ------------------------------------------------------------------ */
/* 3D convolution (FP64, channels-last, zero padding), for the U-Net   */
/* ------------------------------------------------------------------ */
/* x[v][cin], w[cout][taps][cin], y[v][cout]; taps = 27 (ksize 3) or 1 */
static void tap_xyz(int tap, int ksize, int *dz, int *dy, int *dx) {
    if (ksize == 1) {
        *dz = *dy = *dx = 0;
    } else {
        *dz = tap / 9 - 1;
        *dy = (tap / 3) % 3 - 1;
        *dx = tap % 3 - 1;
    }
}

void lgdo_conv3d_forward(const double *x, long n, int cin, int ksize, const double *w, const double *bias, int cout,
                         double *y) {
    const int taps = ksize == 1 ? 1 : 27;
    for (long z = 0; z < n; ++z)
        for (long yy = 0; yy < n; ++yy)
            for (long xx = 0; xx < n; ++xx) {
                const long v = (z * n + yy) * n + xx;
                for (int co = 0; co < cout; ++co) {
                    double s = bias ? bias[co] : 0.0;
                    for (int t = 0; t < taps; ++t) {
                        int dz, dy, dx;
                        tap_xyz(t, ksize, &dz, &dy, &dx);
                        const long zz = z + dz, y2 = yy + dy, x2 = xx + dx;
                        if (zz < 0 || y2 < 0 || x2 < 0 || zz >= n || y2 >= n || x2 >= n) continue;
                        const long u = (zz * n + y2) * n + x2;
                        for (int ci = 0; ci < cin; ++ci) s += w[((long)co * taps + t) * cin + ci] * x[u * cin + ci];
                    }
                    y[v * cout + co] = s;
                }
            }
}

/* dx[v][ci] = sum_{t,co} w[co][t][ci] dy[v - off(t)][co] */
void lgdo_conv3d_dgrad(const double *dy, long n, int cin, int ksize, const double *w, int cout, double *dx) {
    const int taps = ksize == 1 ? 1 : 27;
    const long nv = n * n * n;
    for (long i = 0; i < nv * cin; ++i) dx[i] = 0.0;
    for (long z = 0; z < n; ++z)
        for (long yy = 0; yy < n; ++yy)
            for (long xx = 0; xx < n; ++xx) {
                const long v = (z * n + yy) * n + xx;
                for (int t = 0; t < taps; ++t) {
                    int dz, dyy, dxx;
                    tap_xyz(t, ksize, &dz, &dyy, &dxx);
                    const long zz = z + dz, y2 = yy + dyy, x2 = xx + dxx;
                    if (zz < 0 || y2 < 0 || x2 < 0 || zz >= n || y2 >= n || x2 >= n) continue;
                    const long u = (zz * n + y2) * n + x2;
                    /* y[v] used x[u] with tap t: dx[u] += w * dy[v] */
                    for (int co = 0; co < cout; ++co)
                        for (int ci = 0; ci < cin; ++ci)
                            dx[u * cin + ci] += w[((long)co * taps + t) * cin + ci] * dy[v * cout + co];
                }
            }
}

/* dw[co][t][ci] = sum_v dy[v][co] x[v + off(t)][ci], db[co] = sum_v dy[v][co] */
void lgdo_conv3d_wgrad(const double *x, long n, int cin, int ksize, const double *dy, int cout, double *dw,
                       double *db) {
    const int taps = ksize == 1 ? 1 : 27;
    for (long i = 0; i < (long)cout * taps * cin; ++i) dw[i] = 0.0;
    if (db)
        for (int co = 0; co < cout; ++co) db[co] = 0.0;
    for (long z = 0; z < n; ++z)
        for (long yy = 0; yy < n; ++yy)
            for (long xx = 0; xx < n; ++xx) {
                const long v = (z * n + yy) * n + xx;
                if (db)
                    for (int co = 0; co < cout; ++co) db[co] += dy[v * cout + co];
                for (int t = 0; t < taps; ++t) {
                    int dz, dyy, dxx;
                    tap_xyz(t, ksize, &dz, &dyy, &dxx);
                    const long zz = z + dz, y2 = yy + dyy, x2 = xx + dxx;
                    if (zz < 0 || y2 < 0 || x2 < 0 || zz >= n || y2 >= n || x2 >= n) continue;
                    const long u = (zz * n + y2) * n + x2;
                    for (int co = 0; co < cout; ++co)
                        for (int ci = 0; ci < cin; ++ci)
                            dw[((long)co * taps + t) * cin + ci] += dy[v * cout + co] * x[u * cin + ci];
                }
            }
}
