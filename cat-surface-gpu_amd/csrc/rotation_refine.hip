// Gradient refinement of candidate rotations on the device (no counterpart in
// the reference, which refines its best seed by Nelder-Mead,
// src/cat_surface_gpu/surface_rotation.py:594-699).
//
// Each of m candidates takes `steps` Levenberg-Marquardt steps on a left
// axis-angle perturbation R <- exp([d]x) R, driven by the float64 cost,
// gradient and Gauss-Newton matrix that k_rotation_cost_mats accumulates for
// the float32 copy of the trial matrix: d solves (H + lam diag(H)) d = -g,
// clamped to max_step radians; a trial is kept only if it lowers the cost
// (lam /= 3, floored at 1e-7), else lam *= 4.  Between steps nothing returns
// to the host: a step is the cost launch, its row sums and one small update
// launch.  At the end every candidate's matrix is written with its Euler
// angles (the reference's Rz(Ry Rx) layout, surface_rotation.py:223-313) and
// those angles are costed by the reference's angle-input kernel.
#include "../../include/csg.h"
#include "locate_kernels.h"

using namespace csg;

namespace {

inline hipStream_t S(void *s) { return (hipStream_t)s; }

constexpr int kRefineMax = 256;  // candidates per call (one thread each)

struct RefineState {
    double *rot;    // [m][9] accepted matrices
    double *trial;  // [m][9] matrices being evaluated
    float *trial32; // [m][9] their float32 copies (the cost kernel's input)
    double *cost;   // [m]
    double *grad;   // [m][3]
    double *hess;   // [m][6] xx yy zz xy xz yz
    double *lam;    // [m]
};

// rotation_matrix_np in float64: M = (Rz(g) Ry(-b) Rx(a))ᵀ, row-major.
__device__ void angles_to_matrix(const double *a, double *m)
{
    const double ca = cos(a[0]), sa = sin(a[0]), cb = cos(a[1]), sb = sin(a[1]),
                 cg = cos(a[2]), sg = sin(a[2]);
    // P = Rz(g) Ry(-b) Rx(a) with the column-major tables of the reference
    const double rx[9] = {1, 0, 0, 0, ca, sa, 0, -sa, ca};
    const double ry[9] = {cb, 0, sb, 0, 1, 0, -sb, 0, cb};
    const double rz[9] = {cg, sg, 0, -sg, cg, 0, 0, 0, 1};
    double first[9], res[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            first[r + 3 * c] = ry[r] * rx[3 * c] + ry[r + 3] * rx[1 + 3 * c] +
                               ry[r + 6] * rx[2 + 3 * c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            res[r + 3 * c] = rz[r] * first[3 * c] + rz[r + 3] * first[1 + 3 * c] +
                             rz[r + 6] * first[2 + 3 * c];
    for (int k = 0; k < 9; ++k)
        m[k] = res[k];  // the flat array read row-major is M
}

// Inverse of angles_to_matrix on the cos(b) > 0 branch.
__device__ void matrix_to_angles(const double *m, double *a)
{
    double s = m[2] < -1.0 ? -1.0 : (m[2] > 1.0 ? 1.0 : m[2]);
    a[0] = atan2(m[5], m[8]);
    a[1] = asin(s);
    a[2] = atan2(m[1], m[0]);
}

// exp([d]x) R by Rodrigues' formula, then Gram-Schmidt to keep R orthonormal.
__device__ void rotate_left(const double *d, const double *r, double *out)
{
    const double th2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double th = sqrt(th2);
    const double k[9] = {0.0, -d[2], d[1], d[2], 0.0, -d[0], -d[1], d[0], 0.0};
    double s1, s2;
    if (th < 1e-12) {
        s1 = 1.0;
        s2 = 0.5;
    } else {
        s1 = sin(th) / th;
        s2 = (1.0 - cos(th)) / th2;
    }
    double e[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double kk = k[3 * i] * k[j] + k[3 * i + 1] * k[3 + j] + k[3 * i + 2] * k[6 + j];
            e[3 * i + j] = (i == j ? 1.0 : 0.0) + s1 * k[3 * i + j] + s2 * kk;
        }
    double o[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            o[3 * i + j] = e[3 * i] * r[j] + e[3 * i + 1] * r[3 + j] + e[3 * i + 2] * r[6 + j];
    // re-orthonormalise the rows
    double n0 = rsqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    for (int j = 0; j < 3; ++j)
        o[j] *= n0;
    double p = o[0] * o[3] + o[1] * o[4] + o[2] * o[5];
    for (int j = 0; j < 3; ++j)
        o[3 + j] -= p * o[j];
    double n1 = rsqrt(o[3] * o[3] + o[4] * o[4] + o[5] * o[5]);
    for (int j = 0; j < 3; ++j)
        o[3 + j] *= n1;
    o[6] = o[1] * o[5] - o[2] * o[4];
    o[7] = o[2] * o[3] - o[0] * o[5];
    o[8] = o[0] * o[4] - o[1] * o[3];
    for (int k2 = 0; k2 < 9; ++k2)
        out[k2] = o[k2];
}

// Next trial of candidate j from its accepted state.
__device__ void propose(const RefineState &st, int j, double max_step)
{
    const double *g = st.grad + 3 * j, *h = st.hess + 6 * j;
    double dmax = fmax(fabs(h[0]), fmax(fabs(h[1]), fabs(h[2])));
    double floor_ = 1e-12 * fmax(1.0, dmax);
    double lam = st.lam[j];
    double a00 = h[0] + lam * fmax(h[0], floor_), a11 = h[1] + lam * fmax(h[1], floor_),
           a22 = h[2] + lam * fmax(h[2], floor_), a01 = h[3], a02 = h[4], a12 = h[5];
    // Cramer's rule on the symmetric 3x3 system
    double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    double det = a00 * c00 + a01 * c01 + a02 * c02;
    double d[3];
    if (det != 0.0 && isfinite(det)) {
        double c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
        d[0] = -(c00 * g[0] + c01 * g[1] + c02 * g[2]) / det;
        d[1] = -(c01 * g[0] + c11 * g[1] + c12 * g[2]) / det;
        d[2] = -(c02 * g[0] + c12 * g[1] + c22 * g[2]) / det;
    } else {
        d[0] = -g[0] / (lam * fmax(h[0], floor_) + 1e-300);
        d[1] = -g[1] / (lam * fmax(h[1], floor_) + 1e-300);
        d[2] = -g[2] / (lam * fmax(h[2], floor_) + 1e-300);
    }
    double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (n > max_step) {
        double f = max_step / n;
        d[0] *= f;
        d[1] *= f;
        d[2] *= f;
    }
    double *t = st.trial + 9 * j;
    rotate_left(d, st.rot + 9 * j, t);
    for (int k = 0; k < 9; ++k)
        st.trial32[9 * j + k] = (float)t[k];
}

__global__ void k_refine_init(RefineState st, const double *start, int from_angles, int m,
                              double damping, int *uncovered)
{
    const int j = threadIdx.x;
    if (j == 0)
        *uncovered = 0;
    if (j >= m)
        return;
    double r[9];
    if (from_angles)
        angles_to_matrix(start + 3 * j, r);
    else
        for (int k = 0; k < 9; ++k)
            r[k] = start[9 * j + k];
    for (int k = 0; k < 9; ++k) {
        st.rot[9 * j + k] = r[k];
        st.trial[9 * j + k] = r[k];
        st.trial32[9 * j + k] = (float)r[k];
    }
    st.lam[j] = damping;
    st.cost[j] = INFINITY;
}

// terms: [m][10] sums for the trials (cost, gradient, H); accept or reject,
// then propose the next trial (unless this was the last evaluation).
__global__ void k_refine_step(RefineState st, const double *terms, int m, double max_step,
                              int first, int propose_next)
{
    const int j = threadIdx.x;
    if (j >= m)
        return;
    const double *t = terms + 10 * j;
    if (first || t[0] < st.cost[j]) {
        for (int k = 0; k < 9; ++k)
            st.rot[9 * j + k] = st.trial[9 * j + k];
        st.cost[j] = t[0];
        for (int k = 0; k < 3; ++k)
            st.grad[3 * j + k] = t[1 + k];
        for (int k = 0; k < 6; ++k)
            st.hess[6 * j + k] = t[4 + k];
        if (!first)
            st.lam[j] = fmax(st.lam[j] / 3.0, 1e-7);
    } else {
        st.lam[j] *= 4.0;
    }
    if (propose_next)
        propose(st, j, max_step);
}

__global__ void k_refine_final(RefineState st, int m, double *mats_out, double *angles_out)
{
    const int j = threadIdx.x;
    if (j >= m)
        return;
    for (int k = 0; k < 9; ++k)
        mats_out[9 * j + k] = st.rot[9 * j + k];
    matrix_to_angles(st.rot + 9 * j, angles_out + 3 * j);
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// out[0 .. n) = v (a kernel rather than a memset node, for graph capture)
__global__ void __launch_bounds__(kBlock) k_fill_int(int *__restrict__ out, int v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = v;
}

struct RefineLayout {
    size_t state, terms, partial, hint, total;
};

RefineLayout refine_layout(int ns, int m)
{
    RefineLayout l{};
    l.state = align256(sizeof(double) * (size_t)m * (9 + 9 + 1 + 3 + 6 + 1)) +
              align256(sizeof(float) * (size_t)m * 9);
    l.terms = align256(sizeof(double) * (size_t)m * kRotationGradTerms);
    l.partial = align256(sizeof(double) * (size_t)rotation_cost_blocks(ns) * m *
                         kRotationGradTerms);
    l.hint = align256(sizeof(int) * (size_t)ns * m);
    l.total = l.state + l.terms + l.partial + l.hint;
    return l;
}

}  // namespace

extern "C" {

size_t csg_rotation_refine_workspace_bytes(int ns, int m)
{
    return refine_layout(ns, m).total + 256;
}

int csg_rotation_refine(const float *src_pts, const double *src_vals, int ns,
                        const double *tgt_vals, const int *faces, const float *tri,
                        const float *sph, const int *cell_off, const int *cell_faces, int grid,
                        const double *start, int start_is_angles, int m, int steps,
                        double max_step, double damping, double *mats_out, double *angles_out,
                        double *costs_out, void *ws, size_t ws_bytes, int *uncovered_dev,
                        void *stream)
{
    if (ns < 1 || grid < 1 || m < 1 || m > kRefineMax || steps < 0 || !(max_step > 0.0) ||
        !(damping > 0.0) || start == nullptr || uncovered_dev == nullptr)
        return fail(kInvalidArgument,
                    "invalid rotation refinement arguments (1 <= m <= %d, steps >= 0, "
                    "max_step and damping positive)",
                    kRefineMax);
    const RefineLayout l = refine_layout(ns, m);
    if (ws_bytes < csg_rotation_refine_workspace_bytes(ns, m))
        return fail(kInvalidArgument, "rotation refinement workspace too small");
    hipStream_t s = S(stream);
    char *b = (char *)ws;
    RefineState st{};
    double *d = (double *)b;
    st.rot = d;
    st.trial = d + (size_t)9 * m;
    st.cost = d + (size_t)18 * m;
    st.grad = d + (size_t)19 * m;
    st.hess = d + (size_t)22 * m;
    st.lam = d + (size_t)28 * m;
    st.trial32 = (float *)(b + align256(sizeof(double) * (size_t)m * 29));
    double *terms = (double *)(b + l.state);
    RotationCostArgs a{};
    a.src_pts = src_pts;
    a.src_vals = src_vals;
    a.ns = ns;
    a.tgt_vals = tgt_vals;
    a.faces = faces;
    a.tri = tri;
    a.sph = reinterpret_cast<const float4 *>(sph);
    a.cell_off = cell_off;
    a.cell_faces = cell_faces;
    a.grid = grid;
    a.na = m;
    a.costs = terms;
    a.partial = (double *)(b + l.state + l.terms);
    a.uncovered = uncovered_dev;
    // with face spheres, each step's located faces start the next step's scans
    // (the trials move the points only slightly, so most candidates are then
    // skipped); none for the first
    a.hint = sph ? (int *)(b + l.state + l.terms + l.partial) : nullptr;
    if (a.hint) {
        hipLaunchKernelGGL(k_fill_int, dim3(grid_for((size_t)ns * m)), dim3(kBlock), 0, s,
                           a.hint, -1, (size_t)ns * m);
        CSG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_refine_init, dim3(1), dim3(kRefineMax), 0, s, st, start,
                       start_is_angles, m, damping, uncovered_dev);
    CSG_LAUNCH_CHECK();
    for (int it = 0; it <= steps; ++it) {
        CSG_TRY(launch_rotation_cost_mats(a, st.trial32, true, s));
        hipLaunchKernelGGL(k_refine_step, dim3(1), dim3(kRefineMax), 0, s, st,
                           (const double *)terms, m, max_step, it == 0 ? 1 : 0,
                           it < steps ? 1 : 0);
        CSG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_refine_final, dim3(1), dim3(kRefineMax), 0, s, st, m, mats_out,
                       angles_out);
    CSG_LAUNCH_CHECK();
    // the reference's angle formulation of the final candidates
    a.hint = nullptr;
    a.angles = angles_out;
    a.costs = costs_out;
    a.loc_face = nullptr;
    a.loc_bary = nullptr;
    return launch_rotation_costs(a, s);
}

}  // extern "C"
