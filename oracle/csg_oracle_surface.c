/*
 * CPU oracle, surface side: resampling, heat-kernel smoothing, angle-weighted
 * normals, type-5 curvature, sheet gather, the rotation point locator/cost,
 * and a brute-force closest-triangle search that pins the stencil builder.
 *
 * TEST INFRASTRUCTURE ONLY (see csg_oracle.c).  Paths cited are relative to
 * the reference repository root.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* out[r] = sum_c f32(f64(x[idx[r,c]]) * w[r,c]) accumulated in float32 in corner
 * order.  src/cat_surface_gpu/surface_stencil.py:282-306 */
void oracle_resample_f32(const float *x, const int *idx, const double *w, int n, float *out)
{
#pragma omp parallel for schedule(static) if (n >= 4096)
    for (int r = 0; r < n; ++r)
        for (int d = 0; d < 3; ++d) {
            float acc = 0.0f;
            for (int c = 0; c < 3; ++c)
                acc = acc + (float)((double)x[3 * idx[3 * r + c] + d] * w[3 * r + c]);
            out[3 * r + d] = acc;
        }
}

/* Mean edge length over faces (edges [1,2,0] - [0,1,2]).  surface_stencil.py:349-357 */
double oracle_edge_mean(const float *v, const int *faces, int nf)
{
    double s = 0.0;
    for (int f = 0; f < nf; ++f)
        for (int e = 0; e < 3; ++e) {
            int a = faces[3 * f + e], b = faces[3 * f + (e + 1) % 3];
            double dx = (double)v[3 * b] - v[3 * a];
            double dy = (double)v[3 * b + 1] - v[3 * a + 1];
            double dz = (double)v[3 * b + 2] - v[3 * a + 2];
            s += sqrt(dx * dx + dy * dy + dz * dz);
        }
    return s / (3.0 * nf);
}

/* Heat iterations and sigma.  surface_stencil.py:359-370 */
void oracle_heat_parameters(double edge_mean, double fwhm, int *iterations, double *sigma)
{
    double raw = fwhm * fwhm * 0.541011 / (edge_mean * edge_mean);
    int it = (int)ceil(raw);
    if (it < 1)
        it = 1;
    *iterations = it;
    *sigma = fwhm * 0.7355345 / sqrt((double)it);
}

/* v <- f32((v + sum_nbr v / 3) / (1 + deg / 3)), neighbour sums in float64 over
 * the ascending CSR order.  surface_stencil.py:372-388 */
void oracle_smooth_geometry(const float *v, const int *off, const int *nbr, int n, int iters,
                            float *out)
{
    float *cur = (float *)malloc(sizeof(float) * 3 * n);
    float *nxt = (float *)malloc(sizeof(float) * 3 * n);
    memcpy(cur, v, sizeof(float) * 3 * n);
    for (int it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(static) if (n >= 4096)
        for (int i = 0; i < n; ++i) {
            int deg = off[i + 1] - off[i];
            double den = 1.0 + (double)deg / 3.0;
            for (int d = 0; d < 3; ++d) {
                double s = 0.0;
                for (int k = off[i]; k < off[i + 1]; ++k)
                    s += (double)cur[3 * nbr[k] + d];
                nxt[3 * i + d] = (float)(((double)cur[3 * i + d] + s / 3.0) / den);
            }
        }
        float *t = cur; cur = nxt; nxt = t;
    }
    memcpy(out, cur, sizeof(float) * 3 * n);
    free(cur);
    free(nxt);
}

/* Gaussian edge weights w = exp(-d^2 / (2 sigma^2)); squared_norm selects
 * sum-of-squares (smooth_values_many) over |d|^2 (smooth_values).
 * surface_stencil.py:390-412, :414-476 */
void oracle_heat_weights(const double *geom, const int *off, const int *nbr, int n, double sigma,
                         int squared_norm, double *w)
{
#pragma omp parallel for schedule(static) if (n >= 4096)
    for (int i = 0; i < n; ++i)
        for (int k = off[i]; k < off[i + 1]; ++k) {
            int j = nbr[k];
            double dx = geom[3 * j] - geom[3 * i], dy = geom[3 * j + 1] - geom[3 * i + 1],
                   dz = geom[3 * j + 2] - geom[3 * i + 2];
            double d2;
            if (squared_norm)
                d2 = dx * dx + dy * dy + dz * dz;
            else {
                double d = sqrt(dx * dx + dy * dy + dz * dz);
                d2 = d * d;
            }
            w[k] = exp(-d2 / (2.0 * sigma * sigma));
        }
}

void oracle_smooth_values(const double *vals, const int *off, const int *nbr, const double *w,
                          int n, int iters, double *out)
{
    double *cur = (double *)malloc(sizeof(double) * n);
    double *nxt = (double *)malloc(sizeof(double) * n);
    memcpy(cur, vals, sizeof(double) * n);
    for (int it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(static) if (n >= 4096)
        for (int i = 0; i < n; ++i) {
            double s = 0.0, ws = 0.0;
            for (int k = off[i]; k < off[i + 1]; ++k) {
                s += w[k] * cur[nbr[k]];
                ws += w[k];
            }
            nxt[i] = (cur[i] + s) / (1.0 + ws);
        }
        double *t = cur; cur = nxt; nxt = t;
    }
    memcpy(out, cur, sizeof(double) * n);
    free(cur);
    free(nxt);
}

static inline float norm3f(float x, float y, float z)
{
    return sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
}

/* Angle-weighted vertex normals in float32; contributions summed per vertex in
 * (corner, face) order, as index_add_ applies them.  surface_stencil.py:478-505 */
void oracle_vertex_normals(const float *v, const int *faces, int nf, int n, float *out)
{
    const float eps = FLT_EPSILON;
    float *acc = (float *)calloc((size_t)3 * n, sizeof(float));
    float *fn = (float *)malloc(sizeof(float) * 3 * nf);
    for (int f = 0; f < nf; ++f) {
        const float *p0 = v + 3 * faces[3 * f], *p1 = v + 3 * faces[3 * f + 1],
                    *p2 = v + 3 * faces[3 * f + 2];
        float a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                      a[0] * b[1] - a[1] * b[0]};
        float nr = norm3f(c[0], c[1], c[2]);
        if (nr < eps)
            nr = eps;
        for (int d = 0; d < 3; ++d)
            fn[3 * f + d] = c[d] / nr;
    }
    for (int corner = 0; corner < 3; ++corner)
        for (int f = 0; f < nf; ++f) {
            const float *pc = v + 3 * faces[3 * f + corner];
            const float *pp = v + 3 * faces[3 * f + (corner + 2) % 3];
            const float *pn = v + 3 * faces[3 * f + (corner + 1) % 3];
            float pr[3] = {pp[0] - pc[0], pp[1] - pc[1], pp[2] - pc[2]};
            float fo[3] = {pn[0] - pc[0], pn[1] - pc[1], pn[2] - pc[2]};
            float den = norm3f(pr[0], pr[1], pr[2]) * norm3f(fo[0], fo[1], fo[2]);
            if (den < eps)
                den = eps;
            float dot = pr[0] * fo[0] + pr[1] * fo[1] + pr[2] * fo[2];
            float cosv = dot / den;
            if (cosv < -1.0f) cosv = -1.0f;
            if (cosv > 1.0f) cosv = 1.0f;
            float ang = acosf(cosv);
            int vi = faces[3 * f + corner];
            for (int d = 0; d < 3; ++d)
                acc[3 * vi + d] += fn[3 * f + d] * ang;
        }
    for (int i = 0; i < n; ++i) {
        float nr = norm3f(acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]);
        if (nr < eps)
            nr = eps;
        for (int d = 0; d < 3; ++d)
            out[3 * i + d] = acc[3 * i + d] / nr;
    }
    free(acc);
    free(fn);
}

/* Depth of the surface along the normals of its fwhm-50 smoothed copy:
 * f64(sum_d f32((s - g) * n)).  surface_stencil.py:507-519 */
void oracle_type5_base(const float *geom, const float *smoothed, const float *normals, int n,
                       double *out)
{
    for (int i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int d = 0; d < 3; ++d)
            s = s + (smoothed[3 * i + d] - geom[3 * i + d]) * normals[3 * i + d];
        out[i] = (double)s;
    }
}

/* Type-2 curvature before smoothing.  surface_stencil.py:552-613: per vertex, with
 * float32 normals promoted to float64, the tangent-plane basis from the first
 * neighbour, least-squares shape operator [[a, b], [b, c]] from the projected
 * coordinate and normal differences of all neighbours, and sqrt((k1^2 + k2^2) / 2)
 * of its eigenvalues (0 with fewer than three neighbours). */
static double dot3d(const double *u, const double *v)
{
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

static void tangent(const double *u, const double *nrm, double *out)
{
    double t = dot3d(u, nrm);
    for (int d = 0; d < 3; ++d)
        out[d] = u[d] - t * nrm[d];
}

void oracle_curvature_type2(const float *geom, const float *normals, const int *off,
                            const int *nbr, int n, double *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double nrm[3], p[3], e0[3], e1[3], delta[3];
        for (int d = 0; d < 3; ++d) {
            nrm[d] = normals[3 * i + d];
            p[d] = geom[3 * i + d];
        }
        int first = off[i + 1] > off[i] ? nbr[off[i]] : 0;
        for (int d = 0; d < 3; ++d)
            delta[d] = (double)geom[3 * first + d] - p[d];
        tangent(delta, nrm, e0);
        double len = sqrt(dot3d(e0, e0));
        if (len < 2.220446049250313e-16)
            len = 2.220446049250313e-16;
        for (int d = 0; d < 3; ++d)
            e0[d] /= len;
        /* e1 = (-e0) x n */
        e1[0] = -e0[1] * nrm[2] + e0[2] * nrm[1];
        e1[1] = -e0[2] * nrm[0] + e0[0] * nrm[2];
        e1[2] = -e0[0] * nrm[1] + e0[1] * nrm[0];
        double s1 = 0, s2 = 0, s3 = 0, wx = 0, wy = 0, wxy = 0;
        for (int k = off[i]; k < off[i + 1]; ++k) {
            int j = nbr[k];
            double dc[3], dn[3], tc[3], tn[3];
            for (int d = 0; d < 3; ++d) {
                dc[d] = (double)geom[3 * j + d] - p[d];
                dn[d] = (double)normals[3 * j + d] - nrm[d];
            }
            tangent(dc, nrm, tc);
            tangent(dn, nrm, tn);
            double cx = dot3d(tc, e0), cy = dot3d(tc, e1);
            double nx = dot3d(tn, e0), ny = dot3d(tn, e1);
            s1 += cx * nx;
            s2 += cx * ny + cy * nx;
            s3 += cy * ny;
            wx += cx * cx;
            wy += cy * cy;
            wxy += cx * cy;
        }
        double den = (wx + wy) * (wx * wy - wxy * wxy);
        if (den == 0.0)
            den = 1.0;
        double a = (s3 * wxy * wxy - s2 * wxy * wy + s1 * (wx * wy - wxy * wxy + wy * wy)) / den;
        double b = (s2 * wx * wy - s3 * wx * wxy - s1 * wxy * wy) / den;
        double c = (s1 * wxy * wxy - s2 * wx * wxy + s3 * (wx * wx - wxy * wxy + wx * wy)) / den;
        double tr = a + c, disc = tr * tr - 4.0 * (a * c - b * b);
        double root = disc > 0.0 ? sqrt(disc) : 0.0;
        double k1 = 0.5 * (tr + root), k2 = 0.5 * (tr - root);
        out[i] = off[i + 1] - off[i] > 2 ? sqrt(0.5 * (k1 * k1 + k2 * k2)) : 0.0;
    }
}

/* sheet[p] = sum_c values[idx[p,c]] * w[p,c].  surface_stencil.py:308-317 */
void oracle_sheet_gather(const double *values, const int *idx, const double *w, int m,
                         double *out)
{
    for (int p = 0; p < m; ++p) {
        double s = 0.0;
        for (int c = 0; c < 3; ++c)
            s += values[idx[3 * p + c]] * w[3 * p + c];
        out[p] = s;
    }
}

/* (x - min) / (max - min) */
void oracle_minmax_normalise(double *x, int n)
{
    double lo = x[0], hi = x[0];
    for (int i = 1; i < n; ++i) {
        if (x[i] < lo) lo = x[i];
        if (x[i] > hi) hi = x[i];
    }
    for (int i = 0; i < n; ++i)
        x[i] = (x[i] - lo) / (hi - lo);
}

/* ------------------------------------------------------------------ rotation */

/* Rz (Ry Rx) assembled column-major in float32 from float32 angles.
 * src/cat_surface_gpu/surface_rotation.py:223-292 */
void oracle_rotation_matrix_f32(const double *angles, float *m)
{
    float al = (float)angles[0], be = (float)angles[1], ga = (float)angles[2];
    float ca = cosf(al), sa = sinf(al), cb = cosf(be), sb = sinf(be), cg = cosf(ga),
          sg = sinf(ga);
    float rx[9] = {1, 0, 0, 0, ca, sa, 0, -sa, ca};
    float ry[9] = {cb, 0, sb, 0, 1, 0, -sb, 0, cb};
    float rz[9] = {cg, sg, 0, -sg, cg, 0, 0, 0, 1};
    float first[9], res[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            first[r + 3 * c] = ry[r] * rx[3 * c] + ry[r + 3] * rx[1 + 3 * c] +
                               ry[r + 6] * rx[2 + 3 * c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            res[r + 3 * c] = rz[r] * first[3 * c] + rz[r + 3] * first[1 + 3 * c] +
                             rz[r + 6] * first[2 + 3 * c];
    /* reshape(3, 3) reads the flat column-major array row-major */
    memcpy(m, res, sizeof(res));
}

static inline float dot3(const float *a, const float *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline void cross3(const float *a, const float *b, float *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void seg_dist(const float *q, const float *a, const float *b, float *dist, float *pt)
{
    float dir[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float off[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
    float den = dot3(dir, dir);
    float al = den != 0.0f ? dot3(off, dir) / den : 0.0f;
    if (al < 0.0f) al = 0.0f;
    if (al > 1.0f) al = 1.0f;
    for (int d = 0; d < 3; ++d)
        pt[d] = a[d] + al * dir[d];
    float e[3] = {q[0] - pt[0], q[1] - pt[1], q[2] - pt[2]};
    *dist = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
}

/* Locate q among candidate triangles: plane projection when inside (slack 1e-3),
 * else the nearer of the two edges at the nearest vertex; first minimum wins.
 * Barycentric weights of the chosen point.  surface_rotation.py:354-447, :334-352 */
int oracle_locate(const float *q, const float *tri /* [F][3][3] */, const int *cand, int ncand,
                  float *bary)
{
    float best = INFINITY, bp[3] = {0, 0, 0};
    int bi = -1;
    for (int c = 0; c < ncand; ++c) {
        const float *t = tri + 9 * cand[c];
        const float *p0 = t, *p1 = t + 3, *p2 = t + 6;
        float ea[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        float eb[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float nrm[3];
        cross3(ea, eb, nrm);
        float nsq = dot3(nrm, nrm);
        float off[3] = {p0[0] - q[0], p0[1] - q[1], p0[2] - q[2]};
        float pt = nsq != 0.0f ? dot3(off, nrm) / nsq : 0.0f;
        float proj[3] = {q[0] + pt * nrm[0], q[1] + pt * nrm[1], q[2] + pt * nrm[2]};
        float po[3] = {proj[0] - p0[0], proj[1] - p0[1], proj[2] - p0[2]};
        float xx = dot3(ea, ea), xy = dot3(ea, eb), xv = dot3(ea, po), yy = dot3(eb, eb),
              yv = dot3(eb, po);
        float den = xx * yy - xy * xy;
        float xp = den != 0.0f ? (xv * yy - yv * xy) / den : 0.0f;
        float yp = den != 0.0f ? (yv * xx - xv * xy) / den : 0.0f;
        int inside = den != 0.0f && xp >= -1.0e-3f && xp <= 1.0f + 1.0e-3f && yp >= -1.0e-3f &&
                     yp <= 1.0f + 1.0e-3f && xp + yp >= -1.0e-3f && xp + yp <= 1.0f + 1.0e-3f;
        float dist, cpt[3];
        if (inside) {
            dist = pt * pt * nsq;
            memcpy(cpt, proj, sizeof(cpt));
        } else {
            const float *vs[3] = {p0, p1, p2};
            float dv[3];
            for (int k = 0; k < 3; ++k) {
                float e[3] = {q[0] - vs[k][0], q[1] - vs[k][1], q[2] - vs[k][2]};
                dv[k] = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
            }
            int cv = 0;
            if (dv[1] < dv[cv]) cv = 1;
            if (dv[2] < dv[cv]) cv = 2;
            float d1, d2, pt1[3], pt2[3];
            seg_dist(q, vs[(cv + 2) % 3], vs[cv], &d1, pt1);
            seg_dist(q, vs[cv], vs[(cv + 1) % 3], &d2, pt2);
            if (d1 < d2) {
                dist = d1;
                memcpy(cpt, pt1, sizeof(cpt));
            } else {
                dist = d2;
                memcpy(cpt, pt2, sizeof(cpt));
            }
        }
        if (dist < best) {
            best = dist;
            bi = c;
            memcpy(bp, cpt, sizeof(bp));
        }
    }
    if (bi < 0)
        return -1;
    const float *t = tri + 9 * cand[bi];
    const float *tv[3] = {t, t + 3, t + 6};
    static const int order[3][3] = {{1, 2, 0}, {2, 0, 1}, {0, 1, 2}};
    for (int k = 0; k < 3; ++k) {
        const float *fi = tv[order[k][0]], *se = tv[order[k][1]], *th = tv[order[k][2]];
        float po[3] = {bp[0] - fi[0], bp[1] - fi[1], bp[2] - fi[2]};
        float ho[3] = {se[0] - fi[0], se[1] - fi[1], se[2] - fi[2]};
        float up[3] = {th[0] - fi[0], th[1] - fi[1], th[2] - fi[2]};
        float nn[3], ve[3];
        cross3(ho, up, nn);
        cross3(nn, ho, ve);
        bary[k] = dot3(po, ve) / dot3(up, ve);
    }
    return cand[bi];
}

/* Cell of a point on the [-1, 1]^3 grid.  surface_rotation.py:449-458 */
static inline int cell_of(const float *p, int g)
{
    int c[3];
    for (int d = 0; d < 3; ++d) {
        int k = (int)floorf((p[d] + 1.0f) * ((float)g / 2.0f));
        c[d] = k < 0 ? 0 : (k > g - 1 ? g - 1 : k);
    }
    return c[0] + g * (c[1] + g * c[2]);
}

/* SSD rotation costs for `na` angle triples.  surface_rotation.py:512-546 */
int oracle_rotation_costs(const float *src_pts, const double *src_vals, int ns,
                          const double *tgt_vals, const int *faces, const float *tri,
                          const int *cell_off, const int *cell_faces, int grid,
                          const double *angles, int na, double *costs, int *loc_faces,
                          float *loc_bary)
{
    /* each point's squared difference in parallel, then the sum in point order */
    double *d2 = (double *)malloc(sizeof(double) * (size_t)ns);
    int failed = 0;
    for (int a = 0; a < na && !failed; ++a) {
        float m[9];
        oracle_rotation_matrix_f32(angles + 3 * a, m);
#pragma omp parallel for schedule(static) if (ns >= 4096) reduction(| : failed)
        for (int i = 0; i < ns; ++i) {
            const float *p = src_pts + 3 * i;
            float q[3];
            for (int r = 0; r < 3; ++r)
                q[r] = p[0] * m[3 * r] + p[1] * m[3 * r + 1] + p[2] * m[3 * r + 2];
            int cell = cell_of(q, grid);
            int nc = cell_off[cell + 1] - cell_off[cell];
            float w[3] = {0.0f, 0.0f, 0.0f};
            int f = nc < 1 ? -1 : oracle_locate(q, tri, cell_faces + cell_off[cell], nc, w);
            if (f < 0) {
                failed |= 1;
                d2[i] = 0.0;
                continue;
            }
            double s = tgt_vals[faces[3 * f]] * (double)w[0] +
                       tgt_vals[faces[3 * f + 1]] * (double)w[1] +
                       tgt_vals[faces[3 * f + 2]] * (double)w[2];
            double d = src_vals[i] - s;
            d2[i] = d * d;
            if (loc_faces) {
                loc_faces[(size_t)a * ns + i] = f;
                for (int k = 0; k < 3; ++k)
                    loc_bary[((size_t)a * ns + i) * 3 + k] = w[k];
            }
        }
        double cost = 0.0;
        for (int i = 0; i < ns; ++i)
            cost += d2[i];
        costs[a] = cost;
    }
    free(d2);
    return failed ? -1 : 0;
}

/* ------------------------------------------------------------ closest point */

/* Closest point on triangle (a, b, c) to p with its barycentric weights, by the
 * Voronoi-region method of Ericson, Real-Time Collision Detection, 5.1.5. */
double oracle_closest_on_triangle(const double *p, const double *a, const double *b,
                                  const double *c, double *w)
{
    double ab[3], ac[3], ap[3];
    for (int d = 0; d < 3; ++d) {
        ab[d] = b[d] - a[d];
        ac[d] = c[d] - a[d];
        ap[d] = p[d] - a[d];
    }
#define DOT(u, v) ((u)[0] * (v)[0] + (u)[1] * (v)[1] + (u)[2] * (v)[2])
    double d1 = DOT(ab, ap), d2 = DOT(ac, ap);
    double v, t;
    if (d1 <= 0.0 && d2 <= 0.0) {
        w[0] = 1; w[1] = 0; w[2] = 0;
    } else {
        double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
        double d3 = DOT(ab, bp), d4 = DOT(ac, bp);
        if (d3 >= 0.0 && d4 <= d3) {
            w[0] = 0; w[1] = 1; w[2] = 0;
        } else {
            double vc = d1 * d4 - d3 * d2;
            if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
                v = d1 / (d1 - d3);
                w[0] = 1 - v; w[1] = v; w[2] = 0;
            } else {
                double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
                double d5 = DOT(ab, cp), d6 = DOT(ac, cp);
                if (d6 >= 0.0 && d5 <= d6) {
                    w[0] = 0; w[1] = 0; w[2] = 1;
                } else {
                    double vb = d5 * d2 - d1 * d6;
                    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
                        t = d2 / (d2 - d6);
                        w[0] = 1 - t; w[1] = 0; w[2] = t;
                    } else {
                        double va = d3 * d6 - d5 * d4;
                        if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
                            t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
                            w[0] = 0; w[1] = 1 - t; w[2] = t;
                        } else {
                            double den = 1.0 / (va + vb + vc);
                            v = vb * den;
                            t = vc * den;
                            w[0] = 1 - v - t; w[1] = v; w[2] = t;
                        }
                    }
                }
            }
        }
    }
#undef DOT
    double q[3], e = 0.0;
    for (int d = 0; d < 3; ++d) {
        q[d] = w[0] * a[d] + w[1] * b[d] + w[2] * c[d];
        e += (p[d] - q[d]) * (p[d] - q[d]);
    }
    return e;
}

/* Brute force: closest triangle of the mesh to each query (first minimum). */
void oracle_closest_triangles(const double *verts, const int *faces, int nf, const double *q,
                              int nq, int *face_out, double *w_out)
{
    for (int i = 0; i < nq; ++i) {
        double best = INFINITY, bw[3] = {0, 0, 0};
        int bf = -1;
        for (int f = 0; f < nf; ++f) {
            double w[3];
            double d = oracle_closest_on_triangle(q + 3 * i, verts + 3 * faces[3 * f],
                                                  verts + 3 * faces[3 * f + 1],
                                                  verts + 3 * faces[3 * f + 2], w);
            if (d < best) {
                best = d;
                bf = f;
                memcpy(bw, w, sizeof(bw));
            }
        }
        face_out[i] = bf;
        memcpy(w_out + 3 * i, bw, sizeof(bw));
    }
}
