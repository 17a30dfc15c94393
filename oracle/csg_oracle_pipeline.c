/* CPU oracle, part 3: the pieces the end-to-end CAT_SurfWarp chain needs
 * beyond the operators of csg_oracle.c / csg_oracle_surface.c.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing on the product path links this file; only
 * tests/, the smoke entry point and bench.py's CPU-baseline leg load it.
 *
 *  - oracle_closest_triangles_grid: the exact closest triangle (first minimum
 *    by face index, as the brute-force oracle_closest_triangles) through a
 *    uniform grid searched in growing shells, for full-size stencils.
 *  - oracle_greedy_colours: the greedy vertex colouring in vertex order
 *    (src/cat_surface_gpu/surface_stencil.py:67-90).
 *  - oracle_depth_potential: the rotation feature's depth-potential system
 *    and its solves, restated from src/cat_surface_gpu/rotation_feature.py:
 *    stable normals (:37-75), mixed Voronoi areas (:78-127), system entries
 *    (:130-174), matrix-vector product (:177-190), colour-ordered SOR
 *    (:253-317) and Jacobi-preconditioned CG (:193-250).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS64 2.220446049250313e-16

double oracle_closest_on_triangle(const double *p, const double *a, const double *b,
                                  const double *c, double *w);

/* Threads the oracle's parallel loops use (1 without OpenMP). */
int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------ closest triangles */

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

int oracle_closest_triangles_grid(const double *verts, int nv, const int *faces, int nf,
                                  const double *q, int nq, int *face_out, double *w_out)
{
    if (nf < 1 || nv < 1)
        return -1;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) {
                const double x = verts[3 * faces[3 * f + k] + d];
                lo[d] = x < lo[d] ? x : lo[d];
                hi[d] = x > hi[d] ? x : hi[d];
            }
    int G = (int)(0.5 * sqrt((double)nf));
    G = clampi(G, 4, 256);
    double cell[3];
    for (int d = 0; d < 3; ++d)
        cell[d] = hi[d] > lo[d] ? (hi[d] - lo[d]) / G : 1.0;
    const size_t ncell = (size_t)G * G * G;
    int *box = malloc(sizeof(int) * 6 * (size_t)nf);
    int *count = calloc(ncell + 1, sizeof(int));
    if (!box || !count) {
        free(box);
        free(count);
        return -2;
    }
    for (int f = 0; f < nf; ++f) {
        for (int d = 0; d < 3; ++d) {
            double mn = INFINITY, mx = -INFINITY;
            for (int k = 0; k < 3; ++k) {
                const double x = verts[3 * faces[3 * f + k] + d];
                mn = x < mn ? x : mn;
                mx = x > mx ? x : mx;
            }
            box[6 * f + d] = clampi((int)floor((mn - lo[d]) / cell[d]), 0, G - 1);
            box[6 * f + 3 + d] = clampi((int)floor((mx - lo[d]) / cell[d]), 0, G - 1);
        }
        for (int z = box[6 * f + 2]; z <= box[6 * f + 5]; ++z)
            for (int y = box[6 * f + 1]; y <= box[6 * f + 4]; ++y)
                for (int x = box[6 * f + 0]; x <= box[6 * f + 3]; ++x)
                    count[((size_t)z * G + y) * G + x + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c)
        count[c + 1] += count[c];
    int *list = malloc(sizeof(int) * (size_t)(count[ncell] > 0 ? count[ncell] : 1));
    int *fill = malloc(sizeof(int) * ncell);
    int *stamp = malloc(sizeof(int) * (size_t)nf);
    if (!list || !fill || !stamp) {
        free(box);
        free(count);
        free(list);
        free(fill);
        free(stamp);
        return -2;
    }
    memcpy(fill, count, sizeof(int) * ncell);
    for (int f = 0; f < nf; ++f)
        for (int z = box[6 * f + 2]; z <= box[6 * f + 5]; ++z)
            for (int y = box[6 * f + 1]; y <= box[6 * f + 4]; ++y)
                for (int x = box[6 * f + 0]; x <= box[6 * f + 3]; ++x)
                    list[fill[((size_t)z * G + y) * G + x]++] = f;
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    free(stamp);
    stamp = malloc(sizeof(int) * (size_t)nf * (size_t)nthreads);
    if (!stamp) {
        free(box);
        free(count);
        free(list);
        free(fill);
        return -2;
    }
    for (size_t k = 0; k < (size_t)nf * (size_t)nthreads; ++k)
        stamp[k] = -1;
#pragma omp parallel for schedule(dynamic, 256) if (nq >= 4096)
    for (int i = 0; i < nq; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int *seen = stamp + (size_t)tid * (size_t)nf;
        const double *p = q + 3 * (size_t)i;
        int c0[3];
        for (int d = 0; d < 3; ++d)
            c0[d] = clampi((int)floor((p[d] - lo[d]) / cell[d]), 0, G - 1);
        double best = INFINITY, bw[3] = {0, 0, 0};
        int bf = -1;
        for (int r = 0; r <= G; ++r) {
            /* the cells at Chebyshev distance exactly r from c0 */
            for (int z = c0[2] - r; z <= c0[2] + r; ++z) {
                if (z < 0 || z >= G)
                    continue;
                for (int y = c0[1] - r; y <= c0[1] + r; ++y) {
                    if (y < 0 || y >= G)
                        continue;
                    const int edge_zy = (z == c0[2] - r || z == c0[2] + r ||
                                         y == c0[1] - r || y == c0[1] + r);
                    const int step = edge_zy ? 1 : 2 * r;
                    for (int x = c0[0] - r; x <= c0[0] + r; x += (step > 0 ? step : 1)) {
                        if (x < 0 || x >= G)
                            continue;
                        const size_t cid = ((size_t)z * G + y) * G + x;
                        for (int k = count[cid]; k < count[cid + 1]; ++k) {
                            const int f = list[k];
                            if (seen[f] == i)
                                continue;
                            seen[f] = i;
                            double w[3];
                            const double dd = oracle_closest_on_triangle(
                                p, verts + 3 * faces[3 * f], verts + 3 * faces[3 * f + 1],
                                verts + 3 * faces[3 * f + 2], w);
                            if (dd < best || (dd == best && f < bf)) {
                                best = dd;
                                bf = f;
                                memcpy(bw, w, sizeof(bw));
                            }
                        }
                    }
                }
            }
            /* faces not examined lie outside the box of searched cells: stop when
             * that box's nearest outside point is farther than the best */
            double gap = INFINITY;
            int all = 1;
            for (int d = 0; d < 3; ++d) {
                if (c0[d] - r > 0) {
                    const double g = p[d] - (lo[d] + (c0[d] - r) * cell[d]);
                    gap = g < gap ? g : gap;
                    all = 0;
                }
                if (c0[d] + r < G - 1) {
                    const double g = (lo[d] + (c0[d] + r + 1) * cell[d]) - p[d];
                    gap = g < gap ? g : gap;
                    all = 0;
                }
            }
            if (all)
                break;
            if (bf >= 0 && gap > 0.0) {
                const double safe = gap * (1.0 - 1e-12);
                if (best < safe * safe)
                    break;
            }
        }
        face_out[i] = bf;
        memcpy(w_out + 3 * (size_t)i, bw, sizeof(bw));
    }
    free(box);
    free(count);
    free(list);
    free(fill);
    free(stamp);
    return 0;
}

/* ------------------------------------------------------------ colouring */

/* colours[v] = the smallest colour no earlier-coloured neighbour holds, vertices
 * in index order; returns the colour count. */
int oracle_greedy_colours(const int *off, const int *nbr, int n, int *colours)
{
    int ncol = 0;
    unsigned char *used = calloc(1024, 1);
    if (!used)
        return -1;
    for (int v = 0; v < n; ++v)
        colours[v] = -1;
    for (int v = 0; v < n; ++v) {
        for (int k = off[v]; k < off[v + 1]; ++k) {
            const int c = colours[nbr[k]];
            if (c >= 0 && c < 1024)
                used[c] = 1;
        }
        int c = 0;
        while (c < 1024 && used[c])
            ++c;
        colours[v] = c;
        ncol = c + 1 > ncol ? c + 1 : ncol;
        for (int k = off[v]; k < off[v + 1]; ++k) {
            const int u = colours[nbr[k]];
            if (u >= 0 && u < 1024)
                used[u] = 0;
        }
    }
    free(used);
    return ncol;
}

/* ------------------------------------------------------------ depth potential */

static void sub3(const double *a, const double *b, double *o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

static double dot3d(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void cross3d(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static double norm3d(const double *a) { return sqrt(dot3d(a, a)); }

static double clamp_eps(double x) { return x < EPS64 ? EPS64 : x; }

/* v1 = p1 - p0, v2 = p2 - p1, v3 = p0 - p2 of face f (rotation_feature.py:24-34) */
static void face_vectors(const double *g, const int *faces, int f, double *v1, double *v2,
                         double *v3)
{
    const double *p0 = g + 3 * faces[3 * f], *p1 = g + 3 * faces[3 * f + 1],
                 *p2 = g + 3 * faces[3 * f + 2];
    sub3(p1, p0, v1);
    sub3(p2, p1, v2);
    sub3(p0, p2, v3);
}

/* rotation_feature.py:37-75; out [n][3] */
void oracle_stable_normals(const double *g, const int *faces, int nf, int n, double *out)
{
    double *tn = malloc(sizeof(double) * 3 * (size_t)nf), *wt = malloc(sizeof(double) * 3 * (size_t)nf);
    for (int f = 0; f < nf; ++f) {
        double v[3][3];
        face_vectors(g, faces, f, v[0], v[1], v[2]);
        for (int k = 0; k < 3; ++k) {
            const double l = clamp_eps(norm3d(v[k]));
            for (int d = 0; d < 3; ++d)
                v[k][d] /= l;
        }
        double c31[3], c23[3], c12[3];
        cross3d(v[2], v[0], c31);
        cross3d(v[1], v[2], c23);
        cross3d(v[0], v[1], c12);
        const double m31 = clamp_eps(norm3d(c31)), m23 = clamp_eps(norm3d(c23)),
                     m12 = clamp_eps(norm3d(c12));
        wt[3 * f + 0] = (1.0 - dot3d(v[2], v[0])) / m31;
        wt[3 * f + 1] = (1.0 - dot3d(v[0], v[1])) / m12;
        wt[3 * f + 2] = (1.0 - dot3d(v[1], v[2])) / m23;
        for (int d = 0; d < 3; ++d)
            tn[3 * f + d] = c12[d] / m12;
    }
    memset(out, 0, sizeof(double) * 3 * (size_t)n);
    for (int k = 0; k < 3; ++k)
        for (int f = 0; f < nf; ++f)
            for (int d = 0; d < 3; ++d)
                out[3 * faces[3 * f + k] + d] += tn[3 * f + d] * wt[3 * f + k];
    for (int i = 0; i < n; ++i) {
        const double l = clamp_eps(norm3d(out + 3 * i));
        for (int d = 0; d < 3; ++d)
            out[3 * i + d] /= l;
    }
    free(tn);
    free(wt);
}

/* rotation_feature.py:78-127; out [n] */
void oracle_mixed_voronoi_areas(const double *g, const int *faces, int nf, int n, double *out)
{
    double *con = malloc(sizeof(double) * 3 * (size_t)nf);
    for (int f = 0; f < nf; ++f) {
        double v1[3], v2[3], v3[3], cr[3];
        face_vectors(g, faces, f, v1, v2, v3);
        cross3d(v1, v2, cr);
        const double area = clamp_eps(norm3d(cr));
        const double d31 = dot3d(v3, v1), d12 = dot3d(v1, v2), d23 = dot3d(v2, v3);
        const int b0 = d31 > 0.0, b1 = d12 > 0.0, b2 = d23 > 0.0, bad = b0 || b1 || b2;
        const double l1 = dot3d(v1, v1), l2 = dot3d(v2, v2), l3 = dot3d(v3, v3);
        const double d13 = dot3d(v1, v3);
        const double w0 = l3 * d12 + l1 * d23, w1 = l2 * d13 + l1 * d23, w2 = l2 * d13 + l3 * d12;
        double c[3];
        if (bad) {
            c[0] = c[1] = c[2] = 0.125 * area;
        } else {
            c[0] = -0.125 * w0 / area;
            c[1] = -0.125 * w1 / area;
            c[2] = -0.125 * w2 / area;
        }
        c[0] += b0 ? 0.125 * area : 0.0;
        c[1] += b1 ? 0.125 * area : 0.0;
        c[2] += b2 ? 0.125 * area : 0.0;
        memcpy(con + 3 * f, c, sizeof(c));
    }
    memset(out, 0, sizeof(double) * (size_t)n);
    for (int k = 0; k < 3; ++k)
        for (int f = 0; f < nf; ++f)
            out[faces[3 * f + k]] += con[3 * f + k];
    for (int i = 0; i < n; ++i)
        if (!isfinite(out[i]) || out[i] == 0.0)
            out[i] = 1.0;
    free(con);
}

/* rotation_feature.py:130-174: the 9 entries of each face (blocks of nf in the
 * reference's concatenation order) and, when alpha != 0, alpha * area on the
 * diagonal.  rows/cols/vals hold 9 nf (+ n) entries. */
static int system_entries(const double *g, const int *faces, int nf, int n, const double *areas,
                          double alpha, int *rows, int *cols, double *vals)
{
    for (int f = 0; f < nf; ++f) {
        double v1[3], v2[3], v3[3], cr[3];
        face_vectors(g, faces, f, v1, v2, v3);
        cross3d(v1, v2, cr);
        const double area = clamp_eps(norm3d(cr));
        const double d23 = dot3d(v2, v3) / area, d13 = dot3d(v1, v3) / area,
                     d12 = dot3d(v1, v2) / area;
        const int i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
        const int r[9] = {i0, i1, i2, i0, i1, i1, i2, i0, i2};
        const int c[9] = {i0, i1, i2, i1, i0, i2, i1, i2, i0};
        const double v[9] = {-0.5 * (d23 + d12), -0.5 * (d13 + d23), -0.5 * (d12 + d13),
                             0.5 * d23,          0.5 * d23,          0.5 * d13,
                             0.5 * d13,          0.5 * d12,          0.5 * d12};
        for (int b = 0; b < 9; ++b) {
            rows[(size_t)b * nf + f] = r[b];
            cols[(size_t)b * nf + f] = c[b];
            vals[(size_t)b * nf + f] = v[b];
        }
    }
    int m = 9 * nf;
    if (alpha != 0.0) {
        for (int i = 0; i < n; ++i) {
            rows[m + i] = i;
            cols[m + i] = i;
            vals[m + i] = alpha * areas[i];
        }
        m += n;
    }
    return m;
}

static void matvec(const int *rows, const int *cols, const double *vals, int m, const double *x,
                   int n, double *out)
{
    memset(out, 0, sizeof(double) * (size_t)n);
    for (int k = 0; k < m; ++k)
        out[rows[k]] += vals[k] * x[cols[k]];
}

static double norm_n(const double *x, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        s += x[i] * x[i];
    return sqrt(s);
}

/* Depth potential of one smoothed surface (float32 geometry [n][3] on the
 * stencil sphere's faces).  solver 0: colour-ordered SOR, max_iter sweeps,
 * relaxation omega, colours[v] in [0, ncolours) (rotation_feature.py:253-317);
 * solver 1: Jacobi-preconditioned CG to relative residual tol or max_iter
 * (:193-250).  Writes x [n] and optionally normals [n][3] and areas [n];
 * returns the iteration count, and the final relative residual in *relres. */
int oracle_depth_potential(const float *pts, const int *faces, int nf, int n, double alpha,
                           int solver, int max_iter, double omega, double tol, const int *colours,
                           int ncolours, double *x, double *normals_out, double *areas_out,
                           double *relres)
{
    double *g = malloc(sizeof(double) * 3 * (size_t)n);
    double *nrm = malloc(sizeof(double) * 3 * (size_t)n);
    double *areas = malloc(sizeof(double) * (size_t)n);
    double *zeros = calloc((size_t)n, sizeof(double));
    const size_t cap = 9 * (size_t)nf + (size_t)n;
    int *rows = malloc(sizeof(int) * cap), *cols = malloc(sizeof(int) * cap);
    double *vals = malloc(sizeof(double) * cap);
    double *lap = malloc(sizeof(double) * 3 * (size_t)n);
    double *tmp = malloc(sizeof(double) * (size_t)n), *tmp2 = malloc(sizeof(double) * (size_t)n);
    double *rhs = malloc(sizeof(double) * (size_t)n), *diag = calloc((size_t)n, sizeof(double));
    for (size_t i = 0; i < 3 * (size_t)n; ++i)
        g[i] = (double)pts[i];
    oracle_stable_normals(g, faces, nf, n, nrm);
    oracle_mixed_voronoi_areas(g, faces, nf, n, areas);
    /* mean curvature from the plain Laplacian (alpha = 0), rotation_feature.py:275-290 */
    int m = system_entries(g, faces, nf, n, zeros, 0.0, rows, cols, vals);
    for (int d = 0; d < 3; ++d) {
        for (int i = 0; i < n; ++i)
            tmp[i] = g[3 * i + d];
        matvec(rows, cols, vals, m, tmp, n, tmp2);
        for (int i = 0; i < n; ++i)
            lap[3 * i + d] = tmp2[i] * 2.0;
    }
    double sum_ma = 0.0, sum_a = 0.0;
    for (int i = 0; i < n; ++i) {
        double mc = 0.5 * dot3d(lap + 3 * i, nrm + 3 * i) / areas[i];
        if (!isfinite(mc))
            mc = 0.0;
        tmp[i] = mc;
    }
    for (int i = 0; i < n; ++i)
        sum_ma += tmp[i] * areas[i];
    for (int i = 0; i < n; ++i)
        sum_a += areas[i];
    const double factor = sum_ma / sum_a;
    for (int i = 0; i < n; ++i)
        rhs[i] = (tmp[i] - factor) * areas[i];
    m = system_entries(g, faces, nf, n, areas, alpha, rows, cols, vals);
    for (int k = 0; k < m; ++k)
        if (rows[k] == cols[k])
            diag[rows[k]] += vals[k];
    for (int i = 0; i < n; ++i)
        diag[i] = clamp_eps(diag[i]);
    const double rhs_norm = clamp_eps(norm_n(rhs, n));
    for (int i = 0; i < n; ++i)
        x[i] = 0.0;
    int iterations = 0;
    if (solver == 0) {
        int *fillc = malloc(sizeof(int) * (size_t)ncolours);
        int *vstart = calloc((size_t)ncolours + 1, sizeof(int));
        int *verts = malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; ++i)
            vstart[colours[i] + 1]++;
        for (int c = 0; c < ncolours; ++c)
            vstart[c + 1] += vstart[c];
        memcpy(fillc, vstart, sizeof(int) * (size_t)ncolours);
        for (int i = 0; i < n; ++i)
            verts[fillc[colours[i]]++] = i;
        /* each row's entries in entry order: a row's sum is the reference's
         * index_add order, and rows of one colour (not adjacent) update in parallel */
        int *rstart = calloc((size_t)n + 1, sizeof(int));
        int *rent = malloc(sizeof(int) * (size_t)m);
        for (int k = 0; k < m; ++k)
            rstart[rows[k] + 1]++;
        for (int i = 0; i < n; ++i)
            rstart[i + 1] += rstart[i];
        int *rfill = malloc(sizeof(int) * (size_t)n);
        memcpy(rfill, rstart, sizeof(int) * (size_t)n);
        for (int k = 0; k < m; ++k)
            rent[rfill[rows[k]]++] = k;
        for (int it = 0; it < max_iter; ++it) {
            for (int c = 0; c < ncolours; ++c) {
#pragma omp parallel for schedule(static) if (n >= 4096)
                for (int j = vstart[c]; j < vstart[c + 1]; ++j) {
                    const int i = verts[j];
                    double s = 0.0;
                    for (int e = rstart[i]; e < rstart[i + 1]; ++e)
                        s += vals[rent[e]] * x[cols[rent[e]]];
                    x[i] += omega * (rhs[i] - s) / diag[i];
                }
            }
        }
        free(rstart);
        free(rent);
        free(rfill);
        iterations = max_iter;
        matvec(rows, cols, vals, m, x, n, tmp);
        for (int i = 0; i < n; ++i)
            tmp[i] = rhs[i] - tmp[i];
        *relres = norm_n(tmp, n) / rhs_norm;
        free(fillc);
        free(vstart);
        free(verts);
    } else {
        double *r = malloc(sizeof(double) * (size_t)n), *z = malloc(sizeof(double) * (size_t)n);
        double *p = malloc(sizeof(double) * (size_t)n), *ap = malloc(sizeof(double) * (size_t)n);
        matvec(rows, cols, vals, m, x, n, tmp);
        for (int i = 0; i < n; ++i)
            r[i] = rhs[i] - tmp[i];
        double rel = norm_n(r, n) / rhs_norm;
        if (rel > tol) {
            double rho = 0.0;
            for (int i = 0; i < n; ++i) {
                z[i] = r[i] / diag[i];
                p[i] = z[i];
            }
            for (int i = 0; i < n; ++i)
                rho += r[i] * z[i];
            for (int it = 1; it <= max_iter; ++it) {
                matvec(rows, cols, vals, m, p, n, ap);
                double den = 0.0;
                for (int i = 0; i < n; ++i)
                    den += p[i] * ap[i];
                const double step = rho / den;
                for (int i = 0; i < n; ++i) {
                    x[i] = x[i] + step * p[i];
                    r[i] = r[i] - step * ap[i];
                }
                rel = norm_n(r, n) / rhs_norm;
                iterations = it;
                if (rel <= tol)
                    break;
                double nrho = 0.0;
                for (int i = 0; i < n; ++i)
                    z[i] = r[i] / diag[i];
                for (int i = 0; i < n; ++i)
                    nrho += r[i] * z[i];
                const double beta = nrho / rho;
                for (int i = 0; i < n; ++i)
                    p[i] = z[i] + beta * p[i];
                rho = nrho;
            }
        }
        *relres = rel;
        free(r);
        free(z);
        free(p);
        free(ap);
    }
    if (normals_out)
        memcpy(normals_out, nrm, sizeof(double) * 3 * (size_t)n);
    if (areas_out)
        memcpy(areas_out, areas, sizeof(double) * (size_t)n);
    free(g);
    free(nrm);
    free(areas);
    free(zeros);
    free(rows);
    free(cols);
    free(vals);
    free(lap);
    free(tmp);
    free(tmp2);
    free(rhs);
    free(diag);
    return iterations;
}
