// C ABI of the DARTEL hot path (declared in include/csg.h).
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/csg.h"
#include "dartel_solve.h"
#include "kernel_timing.h"

namespace csg {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

}  // namespace csg

using namespace csg;

#define CSG_DTYPE_DISPATCH(dtype, ...)                                            \
    do {                                                                          \
        if ((dtype) == 0) {                                                       \
            using T = float;                                                      \
            __VA_ARGS__                                                           \
        } else if ((dtype) == 1) {                                                \
            using T = double;                                                     \
            __VA_ARGS__                                                           \
        } else {                                                                  \
            return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)"); \
        }                                                                         \
    } while (0)

#define CSG_REQUIRE(cond, ...)                       \
    do {                                             \
        if (!(cond))                                 \
            return fail(kInvalidArgument, __VA_ARGS__); \
    } while (0)

#define CSG_GRID_CHECK(nx, ny) \
    CSG_REQUIRE((nx) >= 1 && (ny) >= 1 && (size_t)(nx) * (ny) < (1u << 30), \
                "grid dimensions must be positive, got (%d, %d)", (nx), (ny))

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

template <typename T>
static inline const T *C(const void *p) { return static_cast<const T *>(p); }
template <typename T>
static inline T *M(void *p) { return static_cast<T *>(p); }

extern "C" {

const char *csg_last_error(void) { return g_last_error.c_str(); }

const char *csg_version(void) { return "csg-hip 0.1.0 gfx950"; }

int csg_device_count(int *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (count)
        *count = (e == hipSuccess) ? n : 0;
    if (e != hipSuccess)
        return fail(kRuntimeError, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
    return kOk;
}

int csg_set_fmg_schedule(int fused)
{
    return set_fused_fmg_schedule(fused);
}

int csg_set_relax_tile(int nx, int ny, int dtype, int tile_x, int tile_y)
{
    if (dtype != 0 && dtype != 1)
        return fail(kInvalidArgument, "dtype must be 0 or 1");
    return set_relax_tile_shape(nx, ny, dtype == 1 ? 8 : 4, tile_x, tile_y);
}

int csg_fmg_tiled_levels(int nx, int ny, int nit, int dtype)
{
    if (nx < 1 || ny < 1 || nit < 0 || (dtype != 0 && dtype != 1))
        return fail(kInvalidArgument, "invalid FMG shape");
    return fused_tiled_levels(nx, ny, nit, dtype == 1 ? 8 : 4);
}

int csg_get_relax_tile(int nx, int ny, int dtype, int *tile_x, int *tile_y)
{
    if ((dtype != 0 && dtype != 1) || tile_x == nullptr || tile_y == nullptr)
        return fail(kInvalidArgument, "dtype must be 0 or 1 and both outputs given");
    TileShape t = relax_tile_shape(nx, ny, dtype == 1 ? 8 : 4);
    *tile_x = t.x;
    *tile_y = t.y;
    return kOk;
}

int csg_kernel_timing(int enable)
{
    return timing_enable(enable);
}

int csg_kernel_timing_read(int kernel, double *ms, double *bytes, long long *launches)
{
    return timing_read(kernel, ms, bytes, launches);
}

const char *csg_kernel_timing_name(int kernel)
{
    return timed_kernel_name(kernel);
}

int csg_sample_field(const void *field, int nx, int ny, const void *x, const void *y, size_t n,
                     int dtype, void *out, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_DTYPE_DISPATCH(dtype, return launch_sample_field<T>(C<T>(field), nx, ny, C<T>(x), C<T>(y),
                                                            n, M<T>(out), S(stream)););
}

int csg_composition(const void *a, const void *b, int nx, int ny, int dtype, void *out,
                    void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_DTYPE_DISPATCH(dtype, return launch_compose<T>(C<T>(a), nullptr, C<T>(b), nullptr, nx, ny,
                                                       M<T>(out), nullptr, S(stream)););
}

int csg_composition_jacobian(const void *a, const void *ja, const void *b, const void *jb, int nx,
                             int ny, int dtype, void *map_out, void *jac_out, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(ja && jb && jac_out, "composition_jacobian needs both Jacobians");
    CSG_DTYPE_DISPATCH(dtype, return launch_compose<T>(C<T>(a), C<T>(ja), C<T>(b), C<T>(jb), nx,
                                                       ny, M<T>(map_out), M<T>(jac_out),
                                                       S(stream)););
}

int csg_jacobian_of_displacement(const void *v, double scale, int nx, int ny, int dtype,
                                 void *jac_out, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_DTYPE_DISPATCH(dtype, return launch_expdef_init<T>(C<T>(v), scale, nx, ny, 1, 0, nullptr,
                                                           M<T>(jac_out), 0, 0, S(stream)););
}

int csg_smalldef_jac(const void *v, double scale, int nx, int ny, int dtype, void *map_out,
                     void *jac_out, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_DTYPE_DISPATCH(dtype, return launch_expdef_init<T>(C<T>(v), scale, nx, ny, 1, 0,
                                                           M<T>(map_out), M<T>(jac_out), 0, 0,
                                                           S(stream)););
}

int csg_jac_div_smalldef(const void *jac, const void *v, double scale, int nx, int ny, int dtype,
                         void *out, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_DTYPE_DISPATCH(dtype, return launch_jac_div<T>(C<T>(jac), C<T>(v), scale, nx, ny,
                                                       M<T>(out), S(stream)););
}

size_t csg_expdef_workspace_bytes(int nx, int ny, int dtype)
{
    return (size_t)12 * nx * ny * dtype_size(dtype) + 1024;
}

int csg_expdef(const void *v, int nx, int ny, int dtype, int k, void *map_out, void *jac_out,
               void *ws, size_t ws_bytes, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(k >= 0 && k <= 30, "exponentiation step count k must be in [0, 30], got %d", k);
    CSG_REQUIRE(ws_bytes >= csg_expdef_workspace_bytes(nx, ny, dtype),
                "expdef workspace too small");
    CSG_DTYPE_DISPATCH(dtype, {
        size_t P = (size_t)nx * ny;
        T *m[2] = {M<T>(ws), M<T>(ws) + 2 * P};
        T *j[2] = {M<T>(ws) + 4 * P, M<T>(ws) + 8 * P};
        double scale = 1.0 / (double)(1LL << k);
        int rc = launch_expdef_init<T>(C<T>(v), scale, nx, ny, 1, 0, m[0], j[0], 0, 0, S(stream));
        if (rc)
            return rc;
        int e = 0;
        for (int i = 0; i < k; ++i) {
            rc = launch_compose_self<T>(m[e], jac_out ? j[e] : nullptr, nx, ny, 1, 0, 0,
                                        m[e ^ 1], jac_out ? j[e ^ 1] : nullptr, S(stream));
            if (rc)
                return rc;
            e ^= 1;
        }
        CSG_HIP_CHECK(hipMemcpyAsync(map_out, m[e], 2 * P * sizeof(T), hipMemcpyDeviceToDevice,
                                     S(stream)));
        if (jac_out)
            CSG_HIP_CHECK(hipMemcpyAsync(jac_out, j[e], 4 * P * sizeof(T),
                                         hipMemcpyDeviceToDevice, S(stream)));
        return kOk;
    });
}

size_t csg_objfun_workspace_bytes(int nx, int ny)
{
    return (size_t)objective_blocks(nx, ny) * sizeof(double) + 1024;
}

int csg_initialise_objfun(const void *src, const void *tgt, const void *map, const void *jac,
                          const void *dist, int nx, int ny, int dtype, void *rhs, void *hes,
                          double *objective, void *ws, size_t ws_bytes, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(ws_bytes >= csg_objfun_workspace_bytes(nx, ny), "objfun workspace too small");
    CSG_DTYPE_DISPATCH(dtype, {
        ObjectiveArgs<T> o{};
        o.src = C<T>(src);
        o.tgt = C<T>(tgt);
        o.v = nullptr;
        o.map = C<T>(map);
        o.jac = C<T>(jac);
        o.dist = C<T>(dist);
        o.nx = nx;
        o.ny = ny;
        o.batch = 1;
        o.rhs = M<T>(rhs);
        o.hes = M<T>(hes);
        o.partial = static_cast<double *>(ws);
        int rc = launch_objective<T>(o, S(stream));
        if (rc)
            return rc;
        if (objective)
            return launch_reduce_partials(o.partial, objective_blocks(nx, ny), 1, 0.5, objective,
                                          1, 0, S(stream));
        return kOk;
    });
}

size_t csg_squaring_workspace_bytes(int nx, int ny, int dtype)
{
    // two map/Jacobian slots and two b/A slots
    return (size_t)(12 + 10) * nx * ny * dtype_size(dtype) + 4096;
}

int csg_squaring_update(const void *rhs, const void *hes, const void *map, const void *jac, int k,
                        int save_transformation, int nx, int ny, int dtype, void *b_out,
                        void *a_out, void *map_out, void *jac_out, void *scratch_out, void *ws,
                        size_t ws_bytes, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(k >= 0, "squaring step count k must be non-negative, got %d", k);
    CSG_REQUIRE(ws_bytes >= csg_squaring_workspace_bytes(nx, ny, dtype),
                "squaring workspace too small");
    CSG_DTYPE_DISPATCH(dtype, {
        size_t P = (size_t)nx * ny;
        hipStream_t s = S(stream);
        T *w = M<T>(ws);
        T *maps[2] = {w, w + 2 * P};
        T *jacs[2] = {w + 4 * P, w + 8 * P};
        T *bb[2] = {w + 12 * P, w + 14 * P};
        T *ab[2] = {w + 16 * P, w + 19 * P};
        CSG_HIP_CHECK(hipMemcpyAsync(maps[0], map, 2 * P * sizeof(T), hipMemcpyDeviceToDevice, s));
        CSG_HIP_CHECK(hipMemcpyAsync(jacs[0], jac, 4 * P * sizeof(T), hipMemcpyDeviceToDevice, s));
        int cur = 0, oth = 1, pp = 0;
        const T *bc = C<T>(rhs), *ac = C<T>(hes);
        for (int st = 0; st < k; ++st) {
            bool compose = save_transformation || st < k - 1;
            int rc = launch_squaring<T>(bc, ac, maps[cur], jacs[cur], nx, ny, 1, 0, 0, 0, 0,
                                        bb[pp], ab[pp], compose ? nullptr : maps[oth],
                                        compose ? maps[oth] : nullptr,
                                        compose ? jacs[oth] : nullptr, s);
            if (rc)
                return rc;
            bc = bb[pp];
            ac = ab[pp];
            pp ^= 1;
            if (compose) {
                int t = cur;
                cur = oth;
                oth = t;
            }
        }
        if (save_transformation && cur != 0) {
            CSG_HIP_CHECK(hipMemcpyAsync(maps[0], maps[cur], 2 * P * sizeof(T),
                                         hipMemcpyDeviceToDevice, s));
        }
        CSG_HIP_CHECK(hipMemcpyAsync(b_out, bc, 2 * P * sizeof(T), hipMemcpyDeviceToDevice, s));
        CSG_HIP_CHECK(hipMemcpyAsync(a_out, ac, 3 * P * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (map_out)
            CSG_HIP_CHECK(hipMemcpyAsync(map_out, maps[cur], 2 * P * sizeof(T),
                                         hipMemcpyDeviceToDevice, s));
        if (jac_out)
            CSG_HIP_CHECK(hipMemcpyAsync(jac_out, jacs[cur], 4 * P * sizeof(T),
                                         hipMemcpyDeviceToDevice, s));
        if (scratch_out)
            CSG_HIP_CHECK(hipMemcpyAsync(scratch_out, maps[0], 2 * P * sizeof(T),
                                         hipMemcpyDeviceToDevice, s));
        return kOk;
    });
}

int csg_regularization_operator(const void *v, const double *params, int rtype, int nx, int ny,
                                int dtype, void *out, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(params != nullptr, "regularization parameters must contain 5 values");
    if (rtype < 0 || rtype > 2)
        return fail(kInvalidArgument, "regularization type must be 0, 1, or 2, got %d", rtype);
    CSG_DTYPE_DISPATCH(dtype, return launch_regularization<T>(C<T>(v), rtype, params, nx, ny,
                                                              M<T>(out), S(stream)););
}

int csg_apply_membrane_system(const void *hes, const void *u, double w00, double w01, double w10,
                              int nx, int ny, int dtype, void *out, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    MembraneWeights w{w00, w01, w10};
    CSG_DTYPE_DISPATCH(dtype, return launch_apply_membrane<T>(C<T>(hes), C<T>(u), w, nx, ny, 1, 0,
                                                              0, M<T>(out), nullptr, S(stream)););
}

int csg_relax_membrane(const void *hes, const void *rhs, void *solution, double w00, double w01,
                       double w10, int nit, int nx, int ny, int dtype, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(nit >= 0, "relaxation count must be non-negative, got %d", nit);
    MembraneWeights w{w00, w01, w10};
    CSG_DTYPE_DISPATCH(dtype, return launch_relax<T>(C<T>(hes), C<T>(rhs), M<T>(solution), w, nx,
                                                     ny, 1, 0, 0, nit, S(stream)););
}

int csg_resize_field(const void *in, int channels, int snx, int sny, int tnx, int tny, int dtype,
                     void *out, void *stream)
{
    CSG_GRID_CHECK(snx, sny);
    CSG_GRID_CHECK(tnx, tny);
    CSG_REQUIRE(channels >= 1, "resize needs at least one channel");
    CSG_DTYPE_DISPATCH(dtype, return launch_resize<T>(C<T>(in), channels, snx, sny, tnx, tny, 1,
                                                      0, 0, M<T>(out), 0, S(stream)););
}

size_t csg_fmg2_workspace_bytes(int nx, int ny, int batch, int dtype)
{
    return dtype == 1 ? fmg_workspace_bytes<double>(nx, ny, batch)
                      : fmg_workspace_bytes<float>(nx, ny, batch);
}

int csg_fmg2_membrane(const void *hes, const void *rhs, const double *params, int cycles, int nit,
                      const void *initial, int coarse_init, int nx, int ny, int batch, int dtype,
                      void *out, void *ws, size_t ws_bytes, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(cycles >= 0 && nit >= 0, "FMG cycles and nit must be non-negative");
    CSG_REQUIRE(params != nullptr, "regularization parameters must contain 5 values");
    CSG_DTYPE_DISPATCH(dtype, return fmg_run<T>(C<T>(hes), C<T>(rhs), params, cycles, nit,
                                                C<T>(initial), coarse_init, nx, ny, batch,
                                                M<T>(out), ws, ws_bytes, S(stream)););
}

size_t csg_dartel_step_workspace_bytes(int nx, int ny, int batch, int dtype)
{
    return dtype == 1 ? step_workspace_bytes<double>(nx, ny, batch)
                      : step_workspace_bytes<float>(nx, ny, batch);
}

int csg_dartel_step(const void *src, const void *tgt, const void *v, const void *dist, int k,
                    const double *params, double lmreg, int cycles, int nit, int code,
                    int assume_nonzero, int nx, int ny, int batch, int dtype, void *v_out,
                    double *metrics, void *ws, size_t ws_bytes, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(batch >= 1, "batch must be positive");
    CSG_REQUIRE(k >= 0 && k <= 30, "DARTEL step count k must be in [0, 30], got %d", k);
    CSG_REQUIRE(cycles >= 0 && nit >= 0, "DARTEL cycles and nit must be non-negative");
    CSG_REQUIRE(params != nullptr, "DARTEL regularization parameters must contain 5 values");
    if (code != 0 && code != 1)
        return fail(kUnsupported, "dartel_step supports only code 0 and 1");
    CSG_DTYPE_DISPATCH(dtype, {
        StepArgs<T> a{};
        a.src = C<T>(src);
        a.tgt = C<T>(tgt);
        a.v = C<T>(v);
        a.v_out = M<T>(v_out);
        a.dist = C<T>(dist);
        a.dist_batch_stride = (size_t)nx * ny;
        a.k = k;
        for (int i = 0; i < 5; ++i)
            a.params[i] = params[i];
        a.lmreg = lmreg;
        a.cycles = cycles;
        a.nit = nit;
        a.code = code;
        a.assume_nonzero = assume_nonzero;
        a.nx = nx;
        a.ny = ny;
        a.batch = batch;
        a.metrics = metrics;
        a.metrics_stride = 3;
        return dartel_step_run<T>(a, ws, ws_bytes, S(stream));
    });
}

size_t csg_solve_workspace_bytes(int nx, int ny, int batch, int dtype)
{
    return dtype == 1 ? solve_workspace_bytes<double>(nx, ny, batch)
                      : solve_workspace_bytes<float>(nx, ny, batch);
}

int csg_solve_dartel_maps(const void *src, const void *tgt, int nx, int ny, int batch,
                          int steps_per_job, int loop, double lmreg, int cycles, int nit, int its,
                          int code, int assume_nonzero, int dtype, void *flow, void *velocity,
                          double *metrics, void *ws, size_t ws_bytes, void *stream)
{
    CSG_GRID_CHECK(nx, ny);
    CSG_REQUIRE(batch >= 1 && steps_per_job >= 1 && batch % steps_per_job == 0,
                "batch must be a positive multiple of steps_per_job");
    CSG_REQUIRE(loop >= 1, "loop must be positive, got %d", loop);
    CSG_REQUIRE(cycles >= 0 && nit >= 0 && its >= 0, "cycles, nit, and its must be non-negative");
    if (code != 0 && code != 1)
        return fail(kUnsupported, "dartel_step supports only code 0 and 1");
    CSG_DTYPE_DISPATCH(dtype, {
        SolveArgs<T> a{};
        a.src = C<T>(src);
        a.tgt = C<T>(tgt);
        a.nx = nx;
        a.ny = ny;
        a.batch = batch;
        a.steps_per_job = steps_per_job;
        a.loop = loop;
        a.cycles = cycles;
        a.nit = nit;
        a.its = its;
        a.code = code;
        a.assume_nonzero = assume_nonzero;
        a.lmreg = lmreg;
        a.flow = M<T>(flow);
        a.velocity = M<T>(velocity);
        a.metrics = metrics;
        return solve_run<T>(a, ws, ws_bytes, S(stream));
    });
}

}  // extern "C"
