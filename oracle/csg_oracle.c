/*
 * CPU oracle for the CAT_SurfWarp registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain scalar C, the
 * arithmetic of the reference implementation's CPU torch path so that the
 * HIP kernels can be checked against it.  Only tests/, the smoke entry and
 * the CPU-baseline leg of bench.py may load it; the product path never does.
 *
 * Every function names the reference lines it follows (paths relative to the
 * reference repository root).  Planes are channel-major [C][ny][nx], maps are
 * one-based, all DARTEL arithmetic is double precision unless stated.
 *
 * Pinning: each function is compared with golden vectors written by
 * tests/golden/make_golden.py from the reference's own CPU path, and with
 * closed forms / invariants in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PLANE(nx, ny) ((size_t)(nx) * (size_t)(ny))

/* Mixed boundary: x circular, y reflected (period 2*ny), ny==1 -> 0.
 * src/cat_surface_gpu/dartel_grid.py:190-208 */
static inline void bound(long ix, long iy, int nx, int ny, int *ox, int *oy)
{
    long x = ix % nx;
    if (x < 0)
        x += nx;
    *ox = (int)x;
    if (ny == 1) {
        *oy = 0;
        return;
    }
    long period = 2L * ny;
    long r = iy % period;
    if (r < 0)
        r += period;
    *oy = (int)(r >= ny ? period - r - 1 : r);
}

void oracle_bound(const long *ix, const long *iy, int n, int nx, int ny, int *ox, int *oy)
{
    for (int i = 0; i < n; ++i)
        bound(ix[i], iy[i], nx, ny, &ox[i], &oy[i]);
}

typedef struct {
    int i22, i12, i21, i11; /* flat indices into a plane */
    double dx1, dx2, dy1, dy2;
} Corners;

/* Bilinear corner indices and weights at zero-based (x, y).
 * dartel_grid.py:286-343 */
static inline Corners corners_at(double x, double y, int nx, int ny)
{
    Corners c;
    double fx = floor(x), fy = floor(y);
    long ix = (long)fx, iy = (long)fy;
    c.dx1 = x - (double)ix;
    c.dy1 = y - (double)iy;
    c.dx2 = 1.0 - c.dx1;
    c.dy2 = 1.0 - c.dy1;
    int xx, yy;
    bound(ix, iy, nx, ny, &xx, &yy);
    c.i22 = yy * nx + xx;
    bound(ix + 1, iy, nx, ny, &xx, &yy);
    c.i12 = yy * nx + xx;
    bound(ix, iy + 1, nx, ny, &xx, &yy);
    c.i21 = yy * nx + xx;
    bound(ix + 1, iy + 1, nx, ny, &xx, &yy);
    c.i11 = yy * nx + xx;
    return c;
}

static inline double bilinear(const double *f, const Corners *c)
{
    return (f[c->i22] * c->dx2 + f[c->i12] * c->dx1) * c->dy2 +
           (f[c->i21] * c->dx2 + f[c->i11] * c->dx1) * c->dy1;
}

/* Sample a scalar field at zero-based coordinates.  dartel_grid.py:510-531 */
void oracle_sample_field(const double *field, int nx, int ny, const double *x,
                         const double *y, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        Corners c = corners_at(x[i], y[i], nx, ny);
        out[i] = bilinear(field, &c);
    }
}

/* Coordinate sample with corner unwrapping by the period.
 * dartel_grid.py:407-455 (optimized) and :534-549 */
static inline double wrapped(const double *b, const Corners *c, double period)
{
    double k22 = b[c->i22] - 1.0;
    double k12 = b[c->i12] - 1.0;
    double k21 = b[c->i21] - 1.0;
    double k11 = b[c->i11] - 1.0;
    k12 = k12 - floor((k12 - k22) / period + 0.5) * period;
    k21 = k21 - floor((k21 - k22) / period + 0.5) * period;
    k11 = k11 - floor((k11 - k22) / period + 0.5) * period;
    return (k22 * c->dx2 + k12 * c->dx1) * c->dy2 + (k21 * c->dx2 + k11 * c->dx1) * c->dy1 + 1.0;
}

/* out = b o a.  dartel_grid.py:552-575 */
void oracle_composition(const double *a, const double *b, int nx, int ny, double *out)
{
    size_t P = PLANE(nx, ny);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (size_t p = 0; p < P; ++p) {
        Corners c = corners_at(a[p] - 1.0, a[P + p] - 1.0, nx, ny);
        out[p] = wrapped(b, &c, (double)nx);
        out[P + p] = wrapped(b + P, &c, (double)ny);
    }
}

/* (b o a, (Jb o a) * Ja).  dartel_grid.py:458-507, :578-621 */
void oracle_composition_jacobian(const double *a, const double *ja, const double *b,
                                 const double *jb, int nx, int ny, double *out_map,
                                 double *out_jac)
{
    size_t P = PLANE(nx, ny);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (size_t p = 0; p < P; ++p) {
        Corners c = corners_at(a[p] - 1.0, a[P + p] - 1.0, nx, ny);
        double cx = wrapped(b, &c, (double)nx);
        double cy = wrapped(b + P, &c, (double)ny);
        double jb00 = bilinear(jb, &c);
        double jb01 = bilinear(jb + P, &c);
        double jb10 = bilinear(jb + 2 * P, &c);
        double jb11 = bilinear(jb + 3 * P, &c);
        double a0 = ja[p], a1 = ja[P + p], a2 = ja[2 * P + p], a3 = ja[3 * P + p];
        out_jac[p] = jb00 * a0 + jb10 * a1;
        out_jac[P + p] = jb01 * a0 + jb11 * a1;
        out_jac[2 * P + p] = jb00 * a2 + jb10 * a3;
        out_jac[3 * P + p] = jb01 * a2 + jb11 * a3;
        out_map[p] = cx;
        out_map[P + p] = cy;
    }
}

/* Central-difference Jacobian of scale*v, channels [dxvx+1, dxvy, dyvx, dyvy+1].
 * dartel_grid.py:624-652 */
void oracle_jacobian_of_displacement(const double *v, double scale, int nx, int ny,
                                     double *out)
{
    size_t P = PLANE(nx, ny);
    double s = scale * 0.5;
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (int y = 0; y < ny; ++y) {
        for (int x = 0; x < nx; ++x) {
            int xm, ym, xp, yp, x0, y0, x1, y1;
            bound(x - 1, y, nx, ny, &xm, &ym);
            bound(x + 1, y, nx, ny, &xp, &yp);
            bound(x, y - 1, nx, ny, &x0, &y0);
            bound(x, y + 1, nx, ny, &x1, &y1);
            size_t p = (size_t)y * nx + x;
            size_t im = (size_t)ym * nx + xm, ip = (size_t)yp * nx + xp;
            size_t i0 = (size_t)y0 * nx + x0, i1 = (size_t)y1 * nx + x1;
            out[p] = (v[im] * -1.0 + v[ip]) * s + 1.0;
            out[P + p] = (v[P + im] * -1.0 + v[P + ip]) * s;
            out[2 * P + p] = (v[i0] * -1.0 + v[i1]) * s;
            out[3 * P + p] = (v[P + i0] * -1.0 + v[P + i1]) * s + 1.0;
        }
    }
}

/* id + scale*v and its Jacobian.  dartel_grid.py:655-676 */
void oracle_smalldef_jac(const double *v, double scale, int nx, int ny, double *out_map,
                         double *out_jac)
{
    size_t P = PLANE(nx, ny);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            size_t p = (size_t)y * nx + x;
            out_map[p] = (double)(x + 1) + v[p] * scale;
            out_map[P + p] = (double)(y + 1) + v[P + p] * scale;
        }
    oracle_jacobian_of_displacement(v, scale, nx, ny, out_jac);
}

/* J_old * (I + scale*grad v)^-1.  dartel_grid.py:679-720 */
void oracle_jac_div_smalldef(const double *jac, const double *v, double scale, int nx,
                             int ny, double *out)
{
    size_t P = PLANE(nx, ny);
    double *d = (double *)malloc(sizeof(double) * 4 * P);
    oracle_jacobian_of_displacement(v, scale, nx, ny, d);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (size_t p = 0; p < P; ++p) {
        double d00 = d[p], d10 = d[P + p], d01 = d[2 * P + p], d11 = d[3 * P + p];
        double det = d00 * d11 - d01 * d10;
        double i00 = d11 / det, i10 = -d10 / det, i01 = -d01 / det, i11 = d00 / det;
        double o00 = jac[p], o10 = jac[P + p], o01 = jac[2 * P + p], o11 = jac[3 * P + p];
        out[p] = o00 * i00 + o01 * i10;
        out[P + p] = o10 * i00 + o11 * i10;
        out[2 * P + p] = o00 * i01 + o01 * i11;
        out[3 * P + p] = o10 * i01 + o11 * i11;
    }
    free(d);
}

/* Scaling and squaring.  dartel_grid.py:1619-1666 */
void oracle_expdef(const double *v, int k, int with_jac, int nx, int ny, double *out_map,
                   double *out_jac)
{
    size_t P = PLANE(nx, ny);
    double scale = 1.0 / ldexp(1.0, k);
    double *m = (double *)malloc(sizeof(double) * 2 * P);
    double *m2 = (double *)malloc(sizeof(double) * 2 * P);
    double *j = (double *)malloc(sizeof(double) * 4 * P);
    double *j2 = (double *)malloc(sizeof(double) * 4 * P);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            size_t p = (size_t)y * nx + x;
            m[p] = (double)(x + 1) + v[p] * scale;
            m[P + p] = (double)(y + 1) + v[P + p] * scale;
        }
    if (with_jac) {
        oracle_jacobian_of_displacement(v, scale, nx, ny, j);
        for (int s = 0; s < k; ++s) {
            oracle_composition_jacobian(m, j, m, j, nx, ny, m2, j2);
            double *t = m; m = m2; m2 = t;
            t = j; j = j2; j2 = t;
        }
        memcpy(out_jac, j, sizeof(double) * 4 * P);
    } else {
        for (int s = 0; s < k; ++s) {
            oracle_composition(m, m, nx, ny, m2);
            double *t = m; m = m2; m2 = t;
        }
    }
    memcpy(out_map, m, sizeof(double) * 2 * P);
    free(m); free(m2); free(j); free(j2);
}

/* Objective value, gradient (rhs) and Gauss-Newton Hessian.
 * dartel_grid.py:723-775.  Returns 0.5*sum(d^2*w). */
double oracle_initialise_objfun(const double *src, const double *tgt, const double *map,
                                const double *jac, const double *dist, int nx, int ny,
                                double *rhs, double *hes)
{
    size_t P = PLANE(nx, ny);
    double objective = 0.0;
    double *terms = (double *)malloc(sizeof(double) * P);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (size_t p = 0; p < P; ++p) {
        Corners c = corners_at(map[p] - 1.0, map[P + p] - 1.0, nx, ny);
        double k22 = src[c.i22], k12 = src[c.i12], k21 = src[c.i21], k11 = src[c.i11];
        double sampled = (k22 * c.dx2 + k12 * c.dx1) * c.dy2 + (k21 * c.dx2 + k11 * c.dx1) * c.dy1;
        double diff = sampled - tgt[p];
        double ddx = (k11 - k21) * c.dy1 + (k12 - k22) * c.dy2;
        double ddy = (k11 * c.dx1 + k21 * c.dx2) - (k12 * c.dx1 + k22 * c.dx2);
        double gx = jac[p] * ddx + jac[P + p] * ddy;
        double gy = jac[2 * P + p] * ddx + jac[3 * P + p] * ddy;
        double w = dist ? dist[p] : 1.0;
        hes[p] = gx * gx * w;
        hes[P + p] = gy * gy * w;
        hes[2 * P + p] = gx * gy * w;
        rhs[p] = gx * diff * w;
        rhs[P + p] = gy * diff * w;
        terms[p] = diff * diff * w;
    }
    for (size_t p = 0; p < P; ++p)
        objective += terms[p];
    free(terms);
    return 0.5 * objective;
}

/* One fused squaring increment at every pixel: b += det J^T (b o phi),
 * A += det J^T (A o phi) J.  Writes the b increment.  dartel_grid.py:846-885 */
static void squaring_step(double *b, double *a, const double *map, const double *jac,
                          int nx, int ny, double *increment)
{
    size_t P = PLANE(nx, ny);
    double *b0 = (double *)malloc(sizeof(double) * 2 * P);
    double *a0 = (double *)malloc(sizeof(double) * 3 * P);
    memcpy(b0, b, sizeof(double) * 2 * P);
    memcpy(a0, a, sizeof(double) * 3 * P);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (size_t p = 0; p < P; ++p) {
        Corners c = corners_at(map[p] - 1.0, map[P + p] - 1.0, nx, ny);
        double j00 = jac[p], j10 = jac[P + p], j01 = jac[2 * P + p], j11 = jac[3 * P + p];
        double det = j00 * j11 - j01 * j10;
        double sbx = bilinear(b0, &c), sby = bilinear(b0 + P, &c);
        double s00 = bilinear(a0, &c), s11 = bilinear(a0 + P, &c), s01 = bilinear(a0 + 2 * P, &c);
        double ibx = det * (sbx * j00 + sby * j10);
        double iby = det * (sbx * j01 + sby * j11);
        double t1 = s00 * j00 + s01 * j10;
        double t2 = s01 * j00 + s11 * j10;
        double t3 = s00 * j01 + s01 * j11;
        double t4 = s01 * j01 + s11 * j11;
        b[p] = b0[p] + ibx;
        b[P + p] = b0[P + p] + iby;
        a[p] = a0[p] + det * (t1 * j00 + t2 * j10);
        a[P + p] = a0[P + p] + det * (t3 * j01 + t4 * j11);
        a[2 * P + p] = a0[2 * P + p] + det * (t1 * j01 + t2 * j11);
        increment[p] = ibx;
        increment[P + p] = iby;
    }
    free(b0);
    free(a0);
}

/* k squaring updates with the map/Jacobian composed between steps.  The scratch
 * output is the reference's map_buffers[0] after the loop.  Without saving the
 * transformation that is the input map (k=0), the map of the last step (k odd)
 * or the last b increment (k even >= 2); with saving it is the final map.
 * dartel_grid.py:778-914 */
void oracle_squaring_update(const double *rhs, const double *hes, const double *map,
                            const double *jac, int k, int save, int nx, int ny,
                            double *out_b, double *out_a, double *out_map,
                            double *out_jac, double *out_scratch)
{
    size_t P = PLANE(nx, ny);
    double *maps[2], *jacs[2];
    for (int i = 0; i < 2; ++i) {
        maps[i] = (double *)malloc(sizeof(double) * 2 * P);
        jacs[i] = (double *)malloc(sizeof(double) * 4 * P);
    }
    memcpy(maps[0], map, sizeof(double) * 2 * P);
    memcpy(jacs[0], jac, sizeof(double) * 4 * P);
    memcpy(out_b, rhs, sizeof(double) * 2 * P);
    memcpy(out_a, hes, sizeof(double) * 3 * P);
    int cur = 0, oth = 1;
    for (int s = 0; s < k; ++s) {
        /* increment lands in the other slot's map buffer */
        double *inc = (double *)malloc(sizeof(double) * 2 * P);
        squaring_step(out_b, out_a, maps[cur], jacs[cur], nx, ny, inc);
        memcpy(maps[oth], inc, sizeof(double) * 2 * P);
        free(inc);
        if (save || s < k - 1) {
            oracle_composition_jacobian(maps[cur], jacs[cur], maps[cur], jacs[cur], nx, ny,
                                        maps[oth], jacs[oth]);
            int t = cur; cur = oth; oth = t;
        }
    }
    if (save && cur != 0) {
        memcpy(maps[oth], maps[cur], sizeof(double) * 2 * P);
        memcpy(jacs[oth], jacs[cur], sizeof(double) * 4 * P);
    }
    memcpy(out_map, maps[cur], sizeof(double) * 2 * P);
    memcpy(out_jac, jacs[cur], sizeof(double) * 4 * P);
    if (out_scratch)
        memcpy(out_scratch, maps[0], sizeof(double) * 2 * P);
    for (int i = 0; i < 2; ++i) {
        free(maps[i]);
        free(jacs[i]);
    }
}

static inline double at(const double *f, int x, int y, int nx, int ny)
{
    int xx, yy;
    bound(x, y, nx, ny, &xx, &yy);
    return f[(size_t)yy * nx + xx];
}

/* Regulariser L v for rtype 0 (linear elastic), 1 (membrane), 2 (bending).
 * dartel_grid.py:917-1015 */
void oracle_regularization(const double *v, const double *params, int rtype, int nx,
                           int ny, double *out)
{
    size_t P = PLANE(nx, ny);
    double sx = params[0], sy = params[1], mu = params[2], lam = params[3], id = params[4];
    double sx2 = sx * sx, sy2 = sy * sy;
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (int y = 0; y < ny; ++y) {
        for (int x = 0; x < nx; ++x) {
            size_t p = (size_t)y * nx + x;
            for (int c = 0; c < 2; ++c) {
                const double *f = v + c * P;
                double xm = at(f, x - 1, y, nx, ny), xp = at(f, x + 1, y, nx, ny);
                double ym = at(f, x, y - 1, nx, ny), yp = at(f, x, y + 1, nx, ny);
                double r;
                if (rtype == 1) {
                    double w00 = mu * (2.0 * sx2 + 2.0 * sy2) + id;
                    double w01 = -mu * sy2;
                    double w10 = -mu * sx2;
                    r = w00 * f[p] + w01 * (ym + yp) + w10 * (xm + xp);
                } else if (rtype == 0) {
                    double wx0 = mu * (2.0 * sy2 + 4.0 * sx2) + 2.0 * lam * sx2 + id;
                    double wy0 = mu * (2.0 * sx2 + 4.0 * sy2) + 2.0 * lam * sy2 + id;
                    double wx1 = -(2.0 * mu + lam) * sx2;
                    double wy1 = -(2.0 * mu + lam) * sy2;
                    double wx2 = -mu * sx2;
                    double wy2 = -mu * sy2;
                    double wxy = 0.25 * (lam + mu) * sx * sy;
                    const double *g = v + (1 - c) * P;
                    double cross = at(g, x - 1, y - 1, nx, ny) - at(g, x - 1, y + 1, nx, ny) -
                                   at(g, x + 1, y - 1, nx, ny) + at(g, x + 1, y + 1, nx, ny);
                    if (c == 0)
                        r = wx0 * f[p] + wy2 * (ym + yp) + wx1 * (xm + xp) + wxy * cross;
                    else
                        r = wy0 * f[p] + wy1 * (ym + yp) + wx2 * (xm + xp) + wxy * cross;
                } else {
                    double sx4 = sx2 * sx2, sy4 = sy2 * sy2;
                    double w00 = mu * (6.0 * sx4 + 6.0 * sy4 + 8.0 * sx2 * sy2) + id;
                    double w01 = mu * (-4.0 * sx2 * sy2 - 4.0 * sy4);
                    double w02 = mu * sy4;
                    double w10 = mu * (-4.0 * sx4 - 4.0 * sx2 * sy2);
                    double w11 = mu * (2.0 * sx2 * sy2);
                    double w20 = mu * sx4;
                    r = w00 * f[p] + w01 * (ym + yp) +
                        w02 * (at(f, x, y - 2, nx, ny) + at(f, x, y + 2, nx, ny)) +
                        w10 * (xm + xp) +
                        w11 * (at(f, x - 1, y - 1, nx, ny) + at(f, x + 1, y - 1, nx, ny) +
                               at(f, x - 1, y + 1, nx, ny) + at(f, x + 1, y + 1, nx, ny)) +
                        w20 * (at(f, x - 2, y, nx, ny) + at(f, x + 2, y, nx, ny));
                }
                out[c * P + p] = r;
            }
        }
    }
}

/* (L_membrane + A) u.  dartel_grid.py:1018-1070 */
void oracle_apply_membrane(const double *hes, const double *u, const double *params,
                           int nx, int ny, double *out)
{
    size_t P = PLANE(nx, ny);
    oracle_regularization(u, params, 1, nx, ny, out);
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (size_t p = 0; p < P; ++p) {
        double ux = u[p], uy = u[P + p];
        out[p] = out[p] + hes[p] * ux + hes[2 * P + p] * uy;
        out[P + p] = out[P + p] + hes[2 * P + p] * ux + hes[P + p] * uy;
    }
}

/* Red-black Gauss-Seidel on per-pixel 2x2 blocks, in place, row-major within
 * each colour (the first half-sweep updates (x+y) odd).  For even nx this equals
 * the colour-parallel update; for odd nx it is the sequential order of the
 * reference's scalar loop.  dartel_grid.py:1073-1208 */
void oracle_relax(const double *hes, const double *rhs, const double *params, int nit,
                  int nx, int ny, double *u)
{
    size_t P = PLANE(nx, ny);
    double sx = params[0], sy = params[1], mu = params[2];
    double w00 = mu * (2.0 * sx * sx + 2.0 * sy * sy) + params[4];
    double w01 = -mu * sy * sy;
    double w10 = -mu * sx * sx;
    /* rows of one colour are independent when nx is even or 1 (x wraps onto the
     * other colour); for odd nx the reference's order is sequential */
    const int rows_parallel = (nx % 2 == 0) || nx == 1;
    for (int it = 0; it < 2 * nit; ++it) {
#pragma omp parallel for schedule(static) if (rows_parallel && P >= 8192)
        for (int y = 0; y < ny; ++y) {
            int start = ((it % 2) == (y % 2)) ? 1 : 0;
            for (int x = start; x < nx; x += 2) {
                size_t p = (size_t)y * nx + x;
                double rx = rhs[p] - w01 * (at(u, x, y - 1, nx, ny) + at(u, x, y + 1, nx, ny)) -
                            w10 * (at(u, x - 1, y, nx, ny) + at(u, x + 1, y, nx, ny));
                double ry = rhs[P + p] -
                            w01 * (at(u + P, x, y - 1, nx, ny) + at(u + P, x, y + 1, nx, ny)) -
                            w10 * (at(u + P, x - 1, y, nx, ny) + at(u + P, x + 1, y, nx, ny));
                double dx = hes[p] + w00;
                double dy = hes[P + p] + w00;
                double cr = hes[2 * P + p];
                double det = dx * dy * 1.0000000001 - cr * cr;
                u[p] = (dy * rx - cr * ry) / det;
                u[P + p] = (-cr * rx + dx * ry) / det;
            }
        }
    }
}

/* Quadratic B-spline weight.  dartel_grid.py:229-236 */
static inline double wt2(double d)
{
    double a = fabs(d);
    if (a < 0.5)
        return 0.75 - a * a;
    if (a < 1.5) {
        double o = 1.5 - a;
        return 0.5 * o * o;
    }
    return 0.0;
}

/* Separable 3-tap quadratic B-spline resample (restriction and prolongation).
 * dartel_grid.py:239-305 */
void oracle_resize(const double *in, int channels, int snx, int sny, int tnx, int tny,
                   double *out)
{
    size_t SP = PLANE(snx, sny), TP = PLANE(tnx, tny);
    double *mid = (double *)malloc(sizeof(double) * (size_t)tny * snx);
    for (int c = 0; c < channels; ++c) {
        const double *f = in + c * SP;
#pragma omp parallel for schedule(static) if (SP >= 8192)
        for (int ty = 0; ty < tny; ++ty) {
            double loc = ((double)ty + 0.5) * sny / tny - 0.5;
            long o = (long)floor(loc + 0.5);
            double w = wt2((double)o - loc), wm = wt2((double)o - 1.0 - loc),
                   wp = wt2((double)o + 1.0 - loc);
            int d, ym, y0, yp;
            bound(0, o - 1, snx, sny, &d, &ym);
            bound(0, o, snx, sny, &d, &y0);
            bound(0, o + 1, snx, sny, &d, &yp);
            for (int x = 0; x < snx; ++x)
                mid[(size_t)ty * snx + x] = wm * f[(size_t)ym * snx + x] +
                                            w * f[(size_t)y0 * snx + x] +
                                            wp * f[(size_t)yp * snx + x];
        }
#pragma omp parallel for schedule(static) if (SP >= 8192)
        for (int tx = 0; tx < tnx; ++tx) {
            double loc = ((double)tx + 0.5) * snx / tnx - 0.5;
            long o = (long)floor(loc + 0.5);
            double w = wt2((double)o - loc), wm = wt2((double)o - 1.0 - loc),
                   wp = wt2((double)o + 1.0 - loc);
            int d, xm, x0, xp;
            bound(o - 1, 0, snx, sny, &xm, &d);
            bound(o, 0, snx, sny, &x0, &d);
            bound(o + 1, 0, snx, sny, &xp, &d);
            for (int ty = 0; ty < tny; ++ty) {
                const double *row = mid + (size_t)ty * snx;
                out[c * TP + (size_t)ty * tnx + tx] = wm * row[xm] + w * row[x0] + wp * row[xp];
            }
        }
    }
    free(mid);
}

#define MAX_LEVELS 40

/* Full multigrid for (L_membrane + A) u = b.  dartel_grid.py:1211-1417 */
void oracle_fmg2(const double *hes, const double *rhs, const double *params, int cycles,
                 int nit, const double *initial, int assume_nonzero, int nx, int ny,
                 double *out)
{
    int lnx[MAX_LEVELS], lny[MAX_LEVELS];
    int L = 1;
    lnx[0] = nx;
    lny[0] = ny;
    while (lnx[L - 1] >= 2 || lny[L - 1] >= 2) {
        lnx[L] = (lnx[L - 1] + 1) / 2;
        lny[L] = (lny[L - 1] + 1) / 2;
        ++L;
        if (lnx[L - 1] < 2 && lny[L - 1] < 2)
            break;
    }
    double *A[MAX_LEVELS], *B0[MAX_LEVELS], *B[MAX_LEVELS], *U[MAX_LEVELS], *R[MAX_LEVELS];
    double lp[MAX_LEVELS][5];
    for (int l = 0; l < L; ++l) {
        size_t P = PLANE(lnx[l], lny[l]);
        A[l] = (double *)malloc(sizeof(double) * 3 * P);
        B0[l] = (double *)malloc(sizeof(double) * 2 * P);
        B[l] = (double *)calloc(2 * P, sizeof(double));
        U[l] = (double *)calloc(2 * P, sizeof(double));
        R[l] = (double *)malloc(sizeof(double) * 2 * P);
        lp[l][0] = params[0] * lnx[l] / nx;
        lp[l][1] = params[1] * lny[l] / ny;
        lp[l][2] = params[2];
        lp[l][3] = params[3];
        lp[l][4] = params[4];
    }
    size_t P0 = PLANE(nx, ny);
    memcpy(A[0], hes, sizeof(double) * 3 * P0);
    memcpy(B0[0], rhs, sizeof(double) * 2 * P0);
    memcpy(B[0], rhs, sizeof(double) * 2 * P0);
    for (int l = 1; l < L; ++l) {
        oracle_resize(A[l - 1], 3, lnx[l - 1], lny[l - 1], lnx[l], lny[l], A[l]);
        oracle_resize(B0[l - 1], 2, lnx[l - 1], lny[l - 1], lnx[l], lny[l], B0[l]);
    }
    int any_nonzero = 0;
    if (initial) {
        memcpy(U[0], initial, sizeof(double) * 2 * P0);
        for (int l = 1; l < L; ++l)
            oracle_resize(U[l - 1], 2, lnx[l - 1], lny[l - 1], lnx[l], lny[l], U[l]);
        for (size_t p = 0; p < 2 * P0; ++p)
            if (initial[p] != 0.0) {
                any_nonzero = 1;
                break;
            }
    }
    int coarse_init = initial == NULL ? 1 : (assume_nonzero ? 0 : !any_nonzero);
    int last = L - 1;

#define RELAX(l) oracle_relax(A[l], B[l], lp[l], nit, lnx[l], lny[l], U[l])
    if (coarse_init) {
        /* coarse direct 2x2 solve, then relax.  dartel_grid.py:1301-1326 */
        size_t P = PLANE(lnx[last], lny[last]);
        for (size_t p = 0; p < P; ++p) {
            double dx = A[last][p] + params[4];
            double dy = A[last][P + p] + params[4];
            double cr = A[last][2 * P + p];
            double det = dx * dy * 1.0000000001 - cr * cr;
            double bx = B0[last][p], by = B0[last][P + p];
            U[last][p] = (dy * bx - cr * by) / det;
            U[last][P + p] = (-cr * bx + dx * by) / det;
        }
        RELAX(last);
        for (int level = L - 2; level >= 0; --level) {
            oracle_resize(U[level + 1], 2, lnx[level + 1], lny[level + 1], lnx[level],
                          lny[level], U[level]);
            if (level > 0)
                memcpy(B[level], B0[level], sizeof(double) * 2 * PLANE(lnx[level], lny[level]));
            for (int cyc = 0; cyc < cycles; ++cyc) {
                for (int l = level; l < L - 1; ++l) {
                    RELAX(l);
                    size_t P = PLANE(lnx[l], lny[l]);
                    oracle_apply_membrane(A[l], U[l], lp[l], lnx[l], lny[l], R[l]);
#pragma omp parallel for schedule(static) if (P >= 8192)
                    for (size_t p = 0; p < 2 * P; ++p)
                        R[l][p] = B[l][p] - R[l][p];
                    oracle_resize(R[l], 2, lnx[l], lny[l], lnx[l + 1], lny[l + 1], B[l + 1]);
                    memset(U[l + 1], 0, sizeof(double) * 2 * PLANE(lnx[l + 1], lny[l + 1]));
                }
                RELAX(last);
                for (int l = L - 2; l >= level; --l) {
                    size_t P = PLANE(lnx[l], lny[l]);
                    oracle_resize(U[l + 1], 2, lnx[l + 1], lny[l + 1], lnx[l], lny[l], R[l]);
#pragma omp parallel for schedule(static) if (P >= 8192)
                    for (size_t p = 0; p < 2 * P; ++p)
                        U[l][p] = U[l][p] + R[l][p];
                    RELAX(l);
                }
            }
        }
    } else {
        for (int cyc = 0; cyc < cycles; ++cyc) {
            for (int l = 0; l < L - 1; ++l) {
                RELAX(l);
                size_t P = PLANE(lnx[l], lny[l]);
                oracle_apply_membrane(A[l], U[l], lp[l], lnx[l], lny[l], R[l]);
#pragma omp parallel for schedule(static) if (P >= 8192)
                for (size_t p = 0; p < 2 * P; ++p)
                    R[l][p] = B[l][p] - R[l][p];
                oracle_resize(R[l], 2, lnx[l], lny[l], lnx[l + 1], lny[l + 1], B[l + 1]);
                memset(U[l + 1], 0, sizeof(double) * 2 * PLANE(lnx[l + 1], lny[l + 1]));
            }
            RELAX(last);
            for (int l = L - 2; l >= 0; --l) {
                size_t P = PLANE(lnx[l], lny[l]);
                oracle_resize(U[l + 1], 2, lnx[l + 1], lny[l + 1], lnx[l], lny[l], R[l]);
#pragma omp parallel for schedule(static) if (P >= 8192)
                for (size_t p = 0; p < 2 * P; ++p)
                    U[l][p] = U[l][p] + R[l][p];
                RELAX(l);
            }
        }
    }
#undef RELAX
    memcpy(out, U[0], sizeof(double) * 2 * P0);
    for (int l = 0; l < L; ++l) {
        free(A[l]); free(B0[l]); free(B[l]); free(U[l]); free(R[l]);
    }
}

/* One membrane-regularised Gauss-Newton DARTEL update.  metrics = (objective,
 * 0.5 * sum(Lv.v), ||rhs||).  dartel_grid.py:1420-1616 */
void oracle_dartel_step(const double *src, const double *tgt, const double *v, int k,
                        const double *params, double lmreg, int cycles, int nit, int code,
                        const double *dist, int assume_nonzero, int nx, int ny,
                        double *out_v, double *metrics)
{
    size_t P = PLANE(nx, ny);
    double scale = 1.0 / ldexp(1.0, k);
    double *map = (double *)malloc(sizeof(double) * 2 * P);
    double *jac = (double *)malloc(sizeof(double) * 4 * P);
    double *jac2 = (double *)malloc(sizeof(double) * 4 * P);
    double *rhs = (double *)malloc(sizeof(double) * 2 * P);
    double *hes = (double *)malloc(sizeof(double) * 3 * P);
    double *b = (double *)malloc(sizeof(double) * 2 * P);
    double *a = (double *)malloc(sizeof(double) * 3 * P);
    double *smap = (double *)malloc(sizeof(double) * 2 * P);
    double *sjac = (double *)malloc(sizeof(double) * 4 * P);
    double *fmap = (double *)malloc(sizeof(double) * 2 * P);
    double *fjac = (double *)malloc(sizeof(double) * 4 * P);
    double *scratch = (double *)malloc(sizeof(double) * 2 * P);
    double *reg = (double *)malloc(sizeof(double) * 2 * P);
    double *sol = (double *)malloc(sizeof(double) * 2 * P);

    oracle_expdef(v, k, 1, nx, ny, map, jac);
    oracle_jac_div_smalldef(jac, v, scale, nx, ny, jac2);
    double objective = oracle_initialise_objfun(src, tgt, map, jac2, dist, nx, ny, rhs, hes);
    oracle_smalldef_jac(v, -scale, nx, ny, smap, sjac);
    oracle_squaring_update(rhs, hes, smap, sjac, k, code == 1, nx, ny, b, a, fmap, fjac,
                           scratch);
    if (code == 1) {
        double *rb = (double *)malloc(sizeof(double) * 2 * P);
        double *ra = (double *)malloc(sizeof(double) * 3 * P);
        double *rmap = (double *)malloc(sizeof(double) * 2 * P);
        double *rjac = (double *)malloc(sizeof(double) * 4 * P);
        double *rb2 = (double *)malloc(sizeof(double) * 2 * P);
        double *ra2 = (double *)malloc(sizeof(double) * 3 * P);
        oracle_jac_div_smalldef(fjac, v, -scale, nx, ny, jac2);
        objective += oracle_initialise_objfun(tgt, src, fmap, jac2, dist, nx, ny, rb, ra);
        oracle_smalldef_jac(v, scale, nx, ny, rmap, rjac);
        oracle_squaring_update(rb, ra, rmap, rjac, k, 0, nx, ny, rb2, ra2, smap, sjac,
                               scratch);
        for (size_t p = 0; p < 2 * P; ++p)
            b[p] = b[p] - rb2[p];
        for (size_t p = 0; p < 3 * P; ++p)
            a[p] = a[p] + ra2[p];
        free(rb); free(ra); free(rmap); free(rjac); free(rb2); free(ra2);
    }
    oracle_regularization(v, params, 1, nx, ny, reg);
    double reg_obj = 0.0, norm = 0.0;
#pragma omp parallel for schedule(static) if (P >= 8192)
    for (size_t p = 0; p < 2 * P; ++p)
        b[p] = b[p] * scale + reg[p];
    for (size_t p = 0; p < 2 * P; ++p) {
        reg_obj += reg[p] * v[p];
        norm += b[p] * b[p];
    }
    for (size_t p = 0; p < 3 * P; ++p)
        a[p] = a[p] * scale;
    double sp[5] = {params[0], params[1], params[2], params[3], params[4] + lmreg};
    oracle_fmg2(a, b, sp, cycles, nit, scratch, assume_nonzero, nx, ny, sol);
#pragma omp parallel for schedule(static) collapse(2) if (P >= 8192)
    for (int c = 0; c < 2; ++c)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t p = c * P + (size_t)y * nx + x;
                double w = dist ? dist[(size_t)y * nx] : 1.0;
                out_v[p] = v[p] - sol[p] * w;
            }
    metrics[0] = objective;
    metrics[1] = reg_obj * 0.5;
    metrics[2] = sqrt(norm);
    free(map); free(jac); free(jac2); free(rhs); free(hes); free(b); free(a);
    free(smap); free(sjac); free(fmap); free(fjac); free(scratch); free(reg); free(sol);
}

/* Upstream loop parameters: (1, 1, mu, lambda, lambda/2), mu /= 1.25 every 4.
 * src/cat_surface_gpu/surface_warp.py:76-107 */
void oracle_default_params(int loop, double *out)
{
    double mu = 0.125, lam = 0.0;
    for (int i = 0; i < loop; ++i) {
        out[5 * i + 0] = 1.0;
        out[5 * i + 1] = 1.0;
        out[5 * i + 2] = mu;
        out[5 * i + 3] = lam;
        out[5 * i + 4] = lam / 2.0;
        if ((i + 1) % 4 == 0)
            mu /= 1.25;
        lam /= 5.0;
    }
}

/* S independent solves (velocity reset per step), flow = expdef(v_last, 10).
 * metrics has S*loop*its*3 entries.  surface_warp.py:239-367 */
void oracle_solve_dartel_maps(const double *src, const double *tgt, int steps, int loop,
                              double lmreg, int cycles, int nit, int its, int code,
                              int assume_nonzero, int nx, int ny, double *flow,
                              double *metrics)
{
    size_t P = PLANE(nx, ny);
    double *params = (double *)malloc(sizeof(double) * 5 * loop);
    oracle_default_params(loop, params);
    double *v = (double *)malloc(sizeof(double) * 2 * P);
    double *v2 = (double *)malloc(sizeof(double) * 2 * P);
    int m = 0;
    for (int s = 0; s < steps; ++s) {
        memset(v, 0, sizeof(double) * 2 * P);
        for (int l = 0; l < loop; ++l)
            for (int i = 0; i < its; ++i) {
                oracle_dartel_step(src + s * P, tgt + s * P, v, l, params + 5 * l, lmreg,
                                   cycles, nit, code, NULL, assume_nonzero, nx, ny, v2,
                                   metrics + 3 * m);
                ++m;
                double *t = v; v = v2; v2 = t;
            }
    }
    oracle_expdef(v, 10, 0, nx, ny, flow, NULL);
    free(params); free(v); free(v2);
}

/* ------------------------------------------------------------------------ */
/* Final warp of sphere points through a flow.  surface_warp.py:491-584      */

static inline double wrap_disp(double value)
{
    if (value >= 1.0)
        value = value - floor(value);
    if (value <= -1.0)
        value = value + floor(-value);
    if (value >= 0.5)
        value = value - 1.0;
    if (value <= -0.5)
        value = value + 1.0;
    return value;
}

void oracle_wrap_displacement(const double *in, int n, double *out)
{
    for (int i = 0; i < n; ++i)
        out[i] = wrap_disp(in[i]);
}

/* unit: [n][3] float32 unit-sphere points; flow [2][ny][nx]; out [n][3]. */
void oracle_apply_flow(const float *unit, int n, const double *flow, int nx, int ny,
                       int inverse, double *out)
{
    size_t P = PLANE(nx, ny);
    const double pi = 3.14159265358979323846;
    double *ud = (double *)malloc(sizeof(double) * P);
    double *vd = (double *)malloc(sizeof(double) * P);
    for (int r = 0; r < ny; ++r) {
        double w = sin(((r + 0.5) / (double)ny) * pi);
        for (int c = 0; c < nx; ++c) {
            size_t p = (size_t)r * nx + c;
            ud[p] = wrap_disp((flow[p] - c - 1.0) / (double)nx) * w;
            vd[p] = wrap_disp((flow[P + p] - r - 1.0) / (double)ny) * w;
        }
    }
    for (int i = 0; i < n; ++i) {
        double x = unit[3 * i], y = unit[3 * i + 1], z = unit[3 * i + 2];
        if (z < -1.0) z = -1.0;
        if (z > 1.0) z = 1.0;
        double theta = acos(z);
        double cw = -atan2(x, y);
        if (cw < 0.0)
            cw += 2.0 * pi;
        double u = (cw + pi / 2.0) / (2.0 * pi);
        if (u > 1.0)
            u -= 1.0;
        double v = theta / pi;
        Corners c = corners_at(u * nx - 0.5, v * ny - 0.5, nx, ny);
        double su = bilinear(ud, &c), sv = bilinear(vd, &c);
        if (inverse) {
            su = -su;
            sv = -sv;
        }
        u = u + su;
        v = v + sv;
        if (v < 0.0) {
            u = u + 0.5;
            v = -v;
        }
        if (v > 1.0) {
            u = u + 0.5;
            v = 2.0 - v;
        }
        u = u - floor(u);
        if (v < 0.0) v = 0.0;
        if (v > 1.0) v = 1.0;
        double ph = u * (2.0 * pi), th = v * pi;
        double st = sin(th);
        double wx = st * cos(ph), wy = st * sin(ph), wz = cos(th);
        double nrm = sqrt(wx * wx + wy * wy + wz * wz);
        if (nrm < 2.220446049250313e-16)
            nrm = 2.220446049250313e-16;
        out[3 * i] = wx / nrm;
        out[3 * i + 1] = wy / nrm;
        out[3 * i + 2] = wz / nrm;
    }
    free(ud);
    free(vd);
}

/* Latitude-weighted blend of the x-pole and z-pole solutions.
 * src/cat_surface_gpu/gpu_pipeline.py:248-271 */
void oracle_average_xz(const double *xs, const double *zs, int n, float *out)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double xx = xs[3 * i];
        double zz = zs[3 * i + 2];
        xx = xx < -1.0 ? -1.0 : (xx > 1.0 ? 1.0 : xx);
        zz = zz < -1.0 ? -1.0 : (zz > 1.0 ? 1.0 : zz);
        double px = acos(xx) / pi, pz = acos(zz) / pi;
        double wx = exp(-((2.0 * px - 1.0) * (2.0 * px - 1.0)) / 0.1);
        double wz = exp(-((2.0 * pz - 1.0) * (2.0 * pz - 1.0)) / 0.1);
        if (wx < 1e-19) wx = 1e-19;
        if (wz < 1e-19) wz = 1e-19;
        double t = wx + wz;
        double a[3];
        for (int d = 0; d < 3; ++d)
            a[d] = (wx * xs[3 * i + d] + wz * zs[3 * i + d]) / t;
        double nrm = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        if (nrm < 2.220446049250313e-16)
            nrm = 2.220446049250313e-16;
        for (int d = 0; d < 3; ++d)
            out[3 * i + d] = (float)(a[d] / nrm);
    }
}

/* Float32 unit vectors as torch computes them: |p| = sqrt(fma(z,z,fma(y,y,x*x))),
 * clamped below by FLT_EPSILON.  surface_warp.py:518-522 */
void oracle_normalise_f32(const float *in, int n, float *out)
{
    for (int i = 0; i < n; ++i) {
        float x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2];
        float nrm = sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
        if (nrm < 1.1920928955078125e-07f)
            nrm = 1.1920928955078125e-07f;
        out[3 * i] = x / nrm;
        out[3 * i + 1] = y / nrm;
        out[3 * i + 2] = z / nrm;
    }
}
