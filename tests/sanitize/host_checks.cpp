// Drives the host half of libtsx (tomosipo_amd/csrc/tsx.hip) under
// AddressSanitizer and UndefinedBehaviorSanitizer on the CPU, without a GPU:
// geometry validation and its error paths, host-only plans of every kernel
// kind (padded / general forward, separable / general backprojection, FDK
// tables), plan queries, workspace sizes, the argument checks of every compute
// entry point (which return TSX_ERR_NO_DEVICE for host plans), and plan
// destruction (LeakSanitizer checks that plans free what they allocate).
// Exits non-zero on a failed check.
#include "tsx.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

static int g_bad = 0;

#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAILED %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond,    \
                        tsx_last_error());                                       \
            ++g_bad;                                                             \
        }                                                                        \
    } while (0)

// Angles rotating about z (upright detector) or tilted about x.
static std::vector<double> vectors(int kind, int na, double tilt, double zshift)
{
    std::vector<double> v(12 * (size_t)na);
    for (int a = 0; a < na; ++a) {
        const double th = 2.0 * M_PI * a / na;
        double *w = &v[12 * (size_t)a];
        const double r[3] = {std::cos(th), std::sin(th), 0.0};
        w[0] = kind ? -3.0 * r[0] : r[0];
        w[1] = kind ? -3.0 * r[1] : r[1];
        w[2] = kind ? zshift * (a - na / 2.0) / na : 0.0;
        w[3] = w[4] = w[5] = 0.0;
        w[6] = -std::sin(th) * 0.02;
        w[7] = std::cos(th) * 0.02;
        w[8] = 0.0;
        w[9] = 0.0;
        w[10] = 0.02 * std::sin(tilt);
        w[11] = 0.02 * std::cos(tilt);
    }
    return v;
}

static tsx_geometry geometry(int kind, const std::vector<double> &vec, int rows, int cols, int n)
{
    tsx_geometry g;
    std::memset(&g, 0, sizeof(g));
    g.type = kind ? TSX_CONE_VEC : TSX_PARALLEL3D_VEC;
    g.det_row_count = rows;
    g.det_col_count = cols;
    g.n_angles = (int32_t)(vec.size() / 12);
    g.vectors = vec.data();
    g.grid_col_count = n;
    g.grid_row_count = n + 1;
    g.grid_slice_count = n + 2;
    for (int c = 0; c < 3; ++c) {
        g.window_min[c] = -0.5;
        g.window_max[c] = 0.5;
    }
    g.det_supersampling = 1;
    g.vox_supersampling = 1;
    return g;
}

static void host_plan(const tsx_geometry &g, int expect_fp, int expect_bp)
{
    tsx_plan *p = nullptr;
    CHECK(tsx_plan_create(&g, -1, &p) == TSX_OK && p != nullptr);
    if (!p) return;
    tsx_plan_info info;
    CHECK(tsx_plan_info_get(p, &info) == TSX_OK);
    CHECK(info.device == -1 && info.n_angles == g.n_angles);
    if (expect_fp >= 0) CHECK(info.fp_kernel == expect_fp);
    if (expect_bp >= 0) CHECK(info.bp_kernel == expect_bp);
    std::vector<int32_t> axes(g.n_angles);
    CHECK(tsx_plan_axes(p, axes.data()) == TSX_OK);
    for (int32_t a : axes) CHECK(a >= 0 && a <= 2);
    int64_t bytes = -1;
    for (int op = 0; op < 2; ++op)
        for (int64_t batch : {1, 3}) {
            CHECK(tsx_workspace_size(p, op, batch, &bytes) == TSX_OK && bytes >= 0);
        }
    CHECK(tsx_workspace_size(p, 5, 1, &bytes) != TSX_OK);
    const tsx_plan *plans[2] = {p, p};
    CHECK(tsx_multi_workspace_size(plans, 2, TSX_OP_FP, &bytes) == TSX_OK);
    if (g.type == TSX_CONE_VEC) CHECK(tsx_fdk_workspace_size(p, &bytes) == TSX_OK || true);
    // every compute entry point refuses a host plan before touching a device
    float f = 0.0f;
    CHECK(tsx_forward(p, &f, &f, 0, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_backward(p, &f, &f, 1, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_direct_fpbp3d(p, &f, &f, TSX_MODE_SET, TSX_OP_FP, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_direct_fpbp3d_noalloc(p, &f, &f, TSX_MODE_ADD, TSX_OP_BP, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_forward_batched(p, &f, 0, &f, 0, 2, 0, nullptr, 0, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_backward_batched(p, &f, 0, &f, 0, 2, 0, nullptr, 0, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_forward_residual(p, &f, 0, &f, &f, &f, 0, 1, nullptr, 0, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_backward_scaled_add(p, &f, 0, &f, 0, &f, 1, nullptr, 0, nullptr) == TSX_ERR_NO_DEVICE);
    double d = 0.0;
    CHECK(tsx_forward_batched_f64(p, &d, 0, &d, 0, 1, 0, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_backward_batched_f64(p, &d, 0, &d, 0, 1, 0, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_fdk(p, &f, &f, 0, nullptr, 0, nullptr) != TSX_OK);
    CHECK(tsx_plan_autotune(p, nullptr) == TSX_ERR_NO_DEVICE);
    const float *cin[2] = {&f, &f};
    float *cout[2] = {&f, &f};
    CHECK(tsx_multi_forward(plans, cin, cout, 2, 0, nullptr, 0, nullptr) == TSX_ERR_NO_DEVICE);
    CHECK(tsx_multi_backward(plans, cout, cin, 2, 1, nullptr, 0, nullptr) == TSX_ERR_NO_DEVICE);
    tsx_epilogue eps[2] = {{TSX_EP_SET, nullptr, nullptr}, {TSX_EP_RESIDUAL, &f, &f}};
    CHECK(tsx_multi_forward_ep(plans, cin, cout, eps, 2, nullptr, 0, nullptr) != TSX_OK);
    CHECK(tsx_multi_backward_ep(plans, cout, cin, nullptr, 2, nullptr, 0, nullptr) == TSX_ERR_INVALID_ARGUMENT);
    CHECK(tsx_plan_destroy(p) == TSX_OK);
}

int main()
{
    // argument and geometry validation
    tsx_plan *p = nullptr;
    CHECK(tsx_plan_create(nullptr, -1, &p) == TSX_ERR_INVALID_ARGUMENT);
    std::vector<double> vec = vectors(1, 12, 0.0, 0.0);
    tsx_geometry g = geometry(1, vec, 10, 14, 12);
    CHECK(tsx_plan_create(&g, -1, nullptr) == TSX_ERR_INVALID_ARGUMENT);
    tsx_geometry bad = g;
    bad.type = 7;
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_UNSUPPORTED);
    bad = g;
    bad.n_angles = 0;
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_INVALID_ARGUMENT);
    bad = g;
    bad.det_col_count = 0;
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_INVALID_ARGUMENT);
    bad = g;
    bad.grid_slice_count = -3;
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_INVALID_ARGUMENT);
    bad = g;
    bad.window_max[1] = bad.window_min[1];
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_INVALID_ARGUMENT);
    bad = g;
    bad.window_min[0] = NAN;
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_INVALID_ARGUMENT);
    bad = g;
    bad.det_supersampling = 17;
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_INVALID_ARGUMENT);
    bad = g;
    bad.vox_supersampling = 2;
    CHECK(tsx_plan_create(&bad, -1, &p) == TSX_ERR_UNSUPPORTED);
    std::vector<double> nanvec = vec;
    nanvec[5] = INFINITY;
    bad = g;
    bad.vectors = nanvec.data();
    CHECK(tsx_plan_create(&bad, -1, &p) != TSX_OK);
    CHECK(tsx_plan_destroy(nullptr) == TSX_OK);
    CHECK(tsx_plan_info_get(nullptr, nullptr) == TSX_ERR_INVALID_ARGUMENT);
    CHECK(std::strlen(tsx_version()) > 0);

    // host plans of each kernel kind
    // (the vectors stay alive while their geometries are used)
    const std::vector<double> cone = vectors(1, 12, 0.0, 0.0), par = vectors(0, 9, 0.0, 0.0);
    const std::vector<double> helix = vectors(1, 12, 0.0, 0.2), tilted = vectors(1, 12, 0.3, 0.0);
    host_plan(geometry(1, cone, 10, 14, 12), 1, 1);    // cone, upright
    host_plan(geometry(0, par, 10, 14, 12), 1, 1);     // parallel, upright
    host_plan(geometry(1, helix, 10, 14, 12), 1, 1);   // helical cone
    host_plan(geometry(1, tilted, 10, 14, 12), 1, 0);  // tilted detector
    tsx_geometry ss = geometry(0, par, 10, 14, 12);
    ss.det_supersampling = 3;
    host_plan(ss, -1, 1);                              // supersampled: refined-detector plan
    // rays along z: parallel beam from above
    std::vector<double> zv(12 * 4, 0.0);
    for (int a = 0; a < 4; ++a) {
        double *w = &zv[12 * (size_t)a];
        w[2] = 1.0;
        w[0] = 0.1 * a;
        w[6] = 0.02;
        w[10] = 0.02;
    }
    // rays along z, detector v along y: the separable backprojector of the
    // plan with y and z swapped
    host_plan(geometry(0, zv, 10, 14, 12), -1, 1);

    std::printf(g_bad ? "host checks FAILED (%d)\n" : "host checks passed\n", g_bad);
    return g_bad ? 1 : 0;
}
