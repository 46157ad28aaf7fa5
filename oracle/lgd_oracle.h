/*
 * lgd_oracle.h -- the CPU oracle of liblgdhip (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker / the timed CPU baseline.  No product
 * code path (liblgdhip, HipSimulator) calls it.
 *
 * Dense host arrays, x fastest: index (z*ny + y)*nx + x; LBM populations are
 * 19 such planes, member-major.  `threads` > 1 runs the OpenMP variant
 * (the reference's OpenMPSimulator splits planes the same way,
 * src/libgeodecomp/storage/updatefunctormacros.h:9-73).
 *
 * Parity pinning (tests/test_oracle.py), per function:
 *  - lgdo_jacobi3d_f64: the known-answer neighbour sums of
 *    src/libgeodecomp/storage/test/unit/fixedneighborhoodupdatefunctortest.h
 *    (testCube :304-369: value x + 1000y + 1e6z, edge -1001001, expected
 *    encoded coordinate sums, exact in FP64) times 1.0/7.0; constant field
 *    with equal edge is a fixed point; maximum principle; symmetry under
 *    axis permutations; brute-force pure-Python loop on tiny grids.
 *  - lgdo_jacobi3d_f64_wrap: wrap 0 equals lgdo_jacobi3d_f64 bit for bit;
 *    a cyclic shift of the input along wrapped axes shifts the output the
 *    same way (bit for bit); on a full torus a plane wave cos(2 pi k x / nx)
 *    decays by the closed-form eigenvalue (5 + 2 cos(2 pi k / nx)) / 7 per
 *    sweep and the grid sum is conserved; brute force against a
 *    pure-Python periodic loop on tiny grids.
 *  - lgdo_life2d_u8: blinker period 2, glider translation (1,1) per 4
 *    generations, block still life, the reference example's glider and
 *    diehard seeds (src/examples/gameoflife/main.cpp:70-89; diehard vanishes
 *    after exactly 130 generations on an unbounded plane); brute force
 *    against a numpy evaluation of the 3x3 box sum, dead and alive edges.
 *  - lgdo_life2d_u8_wrap: wrap 0 equals lgdo_life2d_u8; a glider on a 16x16
 *    torus crosses both seams and is back in place after 64 generations;
 *    cyclic shifts along wrapped axes commute with the generations; numpy
 *    brute force with np.roll (the periodic 3x3 box sum) for every wrap.
 *  - lgdo_lbm_d3q19_*: rest state is a fixed point; mass conservation with
 *    bounce-back walls and a moving lid; equilibrium moments sum rules
 *    (sum f_eq = rho, sum e f_eq = j, closed form); mirror symmetry of the
 *    cavity about its x-y plane; FP32 and FP64 versions agree to FP32
 *    rounding.
 *  - lgdo_poisson_*: the residual decays by exactly cos(pi/(n+1)) per
 *    sweep for the smoothest eigenvector source (closed form of the Jacobi
 *    spectrum); integer manufactured solutions with integer faces give an
 *    exact right-hand side and are recovered; the FP32 sweep is the FP64
 *    sweep rounded (compared to tolerance) and a brute-force Python loop on
 *    tiny cubes; the weighted sweep returns e unchanged for omega = 0 and
 *    agrees with an FP64 evaluation of e + omega (J(e) - e) from the plain
 *    sweep to a few ulps.
 *  - lgdo_conv3d_*: a delta kernel reproduces (shifted) inputs, linearity,
 *    and the adjoint identities <conv(x), g> = <x, dgrad(g)> and
 *    <conv_w(x), g> = <w, wgrad(x, g)> (exact up to FP64 rounding), plus a
 *    central finite difference of a scalar loss on a tiny shape.
 */
#ifndef LGD_ORACLE_H
#define LGD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* wall flag bits of the cavity cell (same values as liblgdhip's) */
#define LGDO_WALL_XLO 0x01
#define LGDO_WALL_XHI 0x02
#define LGDO_WALL_YLO 0x04
#define LGDO_WALL_YHI 0x08
#define LGDO_WALL_ZLO 0x10
#define LGDO_WALL_ZHI 0x20
#define LGDO_LID_YHI 0x40

int lgdo_jacobi3d_f64(double *grid, long nx, long ny, long nz, double edge, int steps, int threads);
/* periodic axes: wrap bits 1, 2, 4 wrap x, y, z (Topologies::Torus and the
 * selectively wrapped topologies); wrap 0 is lgdo_jacobi3d_f64 */
int lgdo_jacobi3d_f64_wrap(double *grid, long nx, long ny, long nz, double edge, int steps, int threads, int wrap);
int lgdo_life2d_u8(uint8_t *grid, long nx, long ny, int edge, int steps, int threads);
/* periodic axes: wrap bits 1, 2 wrap x, y; wrap 0 is lgdo_life2d_u8 */
int lgdo_life2d_u8_wrap(uint8_t *grid, long nx, long ny, int edge, int steps, int threads, int wrap);

int lgdo_lbm_direction(int i, int *ex, int *ey, int *ez, int *opp);
void lgdo_lbm_cavity_flags(uint8_t *flags, long nx, long ny, long nz);
int lgdo_lbm_d3q19_double(double *f, const uint8_t *flags, long nx, long ny, long nz, double omega, double lux,
                          double luy, double luz, int steps, int threads);
int lgdo_lbm_d3q19_float(float *f, const uint8_t *flags, long nx, long ny, long nz, float omega, float lux,
                         float luy, float luz, int steps, int threads);
void lgdo_lbm_equilibrium_f64(double *f, const double *rho, const double *jx, const double *jy, const double *jz,
                              long ncells);
void lgdo_lbm_moments_f64(const double *f, long ncells, double *rho, double *jx, double *jy, double *jz);

/* Poisson problems: padded (n+2)^3 blocks, x fastest, faces in u's halo */
void lgdo_poisson_set_faces(double *u, long n, const double face[6]);
void lgdo_poisson_sweep_f32(const float *e, float *e2, const float *r, long n);
void lgdo_poisson_sweep_damped_f32(const float *e, float *e2, const float *r, long n, float omega);
double lgdo_poisson_residual_norm(const double *u, const float *f, long n);
long lgdo_poisson_jacobi_f64(double *u, const float *f, long n, double rel_tol, double reference, long max_sweeps,
                             long check, double *final_norm);

/* 3D convolution, FP64, channels-last x[v][cin], w[cout][taps][cin], zero
 * padding; taps = 27 for ksize 3 (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), 1 for 1 */
void lgdo_conv3d_forward(const double *x, long n, int cin, int ksize, const double *w, const double *bias, int cout,
                         double *y);
void lgdo_conv3d_dgrad(const double *dy, long n, int cin, int ksize, const double *w, int cout, double *dx);
void lgdo_conv3d_wgrad(const double *x, long n, int cin, int ksize, const double *dy, int cout, double *dw,
                       double *db);

int lgdo_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
