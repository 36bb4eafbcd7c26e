// libtsx: Joseph forward projection and its exact-transpose backprojection
// for parallel3d_vec and cone_vec geometries on gfx950.
//
// The C ABI is declared in include/tsx.h. This file holds the host-side plan
// builder (FP64 geometry -> per-angle parameters), the kernels and the entry
// points.
//
// Projector definition (the operator tomosipo obtains from ASTRA's "cuda3d"
// projector, tomosipo/astra.py:80-153): for the ray through detector point
// (j, i) with direction D, in grid coordinates g = (X - min)/h - 1/2,
//   y = l * sum_{n=0}^{N_m-1} bilinear(slice n, Q(n)),  Q(n) = P + ((n - P_m)/D_m) D,
//   l = |h * D| / |D_m|,
// with zero outside the volume; the backprojection is the exact transpose.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tsx.h"
#include "tsx_plan.h"
#include "fp_kernels.h"
#include "bp_kernels.h"
#include "fdk_kernels.h"

using tsx::AngleParams;
using namespace tsx;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define TSX_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(TSX_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

// Every entry point that enqueues device work opens a roctx range named
// after itself (visible with rocprofv3 --marker-trace) and passes through
// the fault injector below.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
};

// Timing: with TSX_TIMING=1 an entry point times the device work it enqueues
// on its stream with HIP events and writes "tsx: <entry point> <ms> ms" to
// stderr; the call then waits for its work (a debugging aid, off by default;
// skipped while the stream is being captured into a graph). Failures of the
// event calls only drop the line.
struct EntryTimer {
    const char *name;
    hipStream_t stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EntryTimer(const char *n, void *s) : name(n), stream((hipStream_t)s)
    {
        const char *e = std::getenv("TSX_TIMING");
        if (!(e && e[0] == '1')) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone ||
            hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
            hipEventRecord(e0, stream) != hipSuccess) {
            release();
            (void)hipGetLastError();
        }
    }
    void release()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        e0 = e1 = nullptr;
    }
    ~EntryTimer()
    {
        if (!e0 || !e1) return;
        float ms = 0.0f;
        if (hipEventRecord(e1, stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
            std::fprintf(stderr, "tsx: %s %.3f ms\n", name, ms);
        else
            (void)hipGetLastError();
        release();
    }
};

// Fault injection for tests of failure handling: with TSX_FAIL_LAUNCH set to
// a comma-separated list of call numbers, those calls of the entry points
// that enqueue device work fail with TSX_ERR_HIP before enqueueing anything.
// Calls are numbered from 1 among the calls made while the variable holds its
// current value; the count restarts whenever the value changes.
int injected_failure(const char *entry)
{
    static std::mutex mu;
    static std::string seen;
    static int64_t count = 0;
    const char *e = std::getenv("TSX_FAIL_LAUNCH");
    const std::string val = e ? e : "";
    std::lock_guard<std::mutex> lk(mu);
    if (val != seen) {
        seen = val;
        count = 0;
    }
    if (val.empty()) return TSX_OK;
    ++count;
    const char *q = val.c_str();
    while (*q) {
        char *end = nullptr;
        const long long k = std::strtoll(q, &end, 10);
        if (end == q) break;
        if (k == count)
            return fail(TSX_ERR_HIP, std::string(entry) + ": injected failure of call " +
                                         std::to_string(count) + " (TSX_FAIL_LAUNCH=" + val + ")");
        q = *end == ',' ? end + 1 : end;
    }
    return TSX_OK;
}

#define TSX_ENTRY(name)                          \
    TraceRange trace_range_(name);               \
    do {                                         \
        int inj_ = injected_failure(name);       \
        if (inj_) return inj_;                   \
    } while (0)

struct tsx_plan {
    int device = -1;
    int32_t kind = 0;
    int32_t n_angles = 0, rows = 0, cols = 0;
    int32_t nx = 0, ny = 0, nz = 0;
    int32_t ss = 1;  // detector supersampling
    std::vector<AngleParams> host;
    // device: the n_angles entries of `host`, then the same entries ordered by
    // stepping axis (x, y, z; stable), each with its index in `host` in `pad`
    // (the angle order of the wave-local separable backprojector)
    AngleParams *dev = nullptr;
    // separable backprojector: slice-range table (bp_range_table), device
    uint32_t *bp_rt = nullptr;
    int64_t bp_rt_stride = 0;  // entries per angle
    int32_t axis_count[3] = {0, 0, 0};
    // padded per-axis layouts for the forward projector (elements per layout)
    int64_t layout_elems[3] = {0, 0, 0};
    bool fp_padded = false;   // padded forward projector (on the whole volume or in slabs)
    bool fp_pair = false;     // ... also with row-paired (8-byte) elements
    // largest byte size of one padded layout or of one backprojection launch's
    // transposed projections (chunk_limit() when the plan was created)
    int64_t chunk_bytes = (int64_t)1 << 32;
    // z-slabs of the forward projector when the whole volume's layouts reach
    // chunk_bytes: plans of the same rays over slices [slab_z0[k], slab_z0[k+1])
    std::vector<tsx_plan *> slabs;
    std::vector<int32_t> slab_z0;
    bool is_slab = false;
    int32_t fp_slab_depth = 0;  // slices per slab (nz: no slabs)
    bool bp_separable = false;  // upright detector, x/y stepping, ss == 1
    int32_t bp_chunk = 0;       // separable backprojector: sorted angles per launch
    // detector v parallel to x (swap 1) or y (2) at every angle: the plan of
    // the same scan with that axis and z swapped, whose detector is upright
    // (separable backprojector; forward layouts with u along their rows)
    tsx_plan *perm = nullptr;
    int32_t swap = 0;
    // detector supersampling ss > 1: the plan of the same scan on the detector
    // refined ss times (ss = 1); its separable backprojector (directly or
    // through its axis swap) takes this plan's backprojection, its padded
    // forward projector this plan's forward projection (bp_refined,
    // fp_refined)
    tsx_plan *fine = nullptr;
    // forward tiles sheared along each angle's shear: the largest offset of a
    // lane across the waves (rows, or columns for column waves), ceil(32 max
    // |shear|) (0: unsheared tiles; also 0 where it would reach half the
    // detector's extent across the waves)
    int32_t fp_shear_rows = 0;
    int32_t fp_shear = -1;  // sheared tiles chosen by tsx_plan_autotune: 1 on, 0 off (-1: on)
    // forward tiles whose waves run along detector columns (every angle's rays
    // keep to one row of the slices along a direction closer to the
    // detector's v than to its u); the shear is then in columns per row
    bool fp_column_waves = false;
    // padded layouts whose rows reach 2^24 bytes: kernels with 32-bit row
    // offset multiplies, marches 3 and 4 on tiles of 8 rows
    bool fp_wide = false;
    std::string bp_reason;      // why bp_separable is off
    // kernel variants chosen by tsx_plan_autotune (-1: library default)
    int32_t fp_variant = -1, bp_bz = -1, fp_by = -1;
    // FDK (cone plans): per-angle quantities and the volume window
    std::vector<FdkAngle> fdk_host;
    FdkAngle *fdk_dev = nullptr;
    bool fdk_upright = false;  // every angle: v parallel to z, normal perpendicular to z
    double vlo[3] = {0, 0, 0}, vh[3] = {0, 0, 0};
};

// ---------------------------------------------------------------------------
// host plan builder (FP64)
// ---------------------------------------------------------------------------
namespace {

struct V3 {
    double v[3];
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};
V3 mk(double a, double b, double c) { return V3{{a, b, c}}; }
V3 operator+(V3 a, V3 b) { return mk(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
V3 operator-(V3 a, V3 b) { return mk(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
V3 operator*(double s, V3 a) { return mk(s * a[0], s * a[1], s * a[2]); }
double dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 cross(V3 a, V3 b)
{
    return mk(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
bool finite3(V3 a) { return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]); }

// Derivatives (dQ1/dj, dQ1/di) of the c1 position at which the rays of an
// angle meet a slice, at the detector's centre (up to a common factor): with
// Q1 = S1 + (n - S0) D1 / D0 (cone, D = P - S) or P1 + (n - P0_m) S1 / S0
// (parallel, D = S), dQ1/dj is proportional to U1 - U0 D1 / D0 and dQ1/di to
// V1 - V0 D1 / D0. The rays keep to one row of the slices along di/dj =
// -dj/di: the forward projector's tiles follow that direction (an upright
// detector has dQ1/dj = 0: waves along its rows, unsheared).
void tile_slopes(const AngleParams &A, int rows, int cols, double *dj, double *di)
{
    const double jc = 0.5 * (cols - 1), ic = 0.5 * (rows - 1);
    double D[3];
    for (int k = 0; k < 3; ++k) {
        const double P = A.P0[k] + jc * A.U[k] + ic * A.V[k];
        D[k] = A.cone ? P - A.S[k] : A.S[k];
    }
    *dj = *di = 0.0;
    if (D[0] == 0.0) return;
    *dj = A.U[1] - A.U[0] * D[1] / D[0];
    *di = A.V[1] - A.V[0] * D[1] / D[0];
}

// Shear of a wave (offset across the wave per lane) from the slopes: waves
// along detector rows take -dj/di rows per column, waves along columns -di/dj
// columns per row; shears under 1/64 (less than half a row across a wave) are
// taken as 0 and shears are held to [-1, 1].
float tile_shear(double dj, double di, bool column)
{
    const double num = column ? -di : -dj, den = column ? dj : di;
    if (!(std::fabs(den) > 0.0)) return 0.0f;
    const double t = num / den;
    if (!(std::fabs(t) >= 1.0 / 64.0)) return 0.0f;
    return (float)std::max(-1.0, std::min(1.0, t));
}

// The per-angle terms of a parallel-beam angle that the general
// backprojector would otherwise form per voxel: the detector position as an
// affine function of the grid point (t0 = (P0 - k) . N / EN, R = k + t0 S -
// P0, j = R . Bj, i = R . Bi), the local map from detector offsets to
// in-slice offsets (rays with slopes r = S / S_m) and the bound of the
// image of a voxel's +-1 square on the detector.
void parallel_map(AngleParams &A)
{
    for (int k = 0; k < 8; ++k) A.par_c[k] = 0.0;
    for (int k = 0; k < 4; ++k) A.par_a[k] = 0.0f;
    A.par_b[0] = A.par_b[1] = -1.0f;
    if (A.cone || A.EN == 0.0 || A.S[0] == 0.0) return;
    const double *B[2] = {A.Bj, A.Bi};
    const double PN = A.P0[0] * A.N[0] + A.P0[1] * A.N[1] + A.P0[2] * A.N[2];
    for (int r = 0; r < 2; ++r) {
        const double SB = A.S[0] * B[r][0] + A.S[1] * B[r][1] + A.S[2] * B[r][2];
        const double PB = A.P0[0] * B[r][0] + A.P0[1] * B[r][1] + A.P0[2] * B[r][2];
        A.par_c[4 * r] = PN * SB * A.inv_EN - PB;
        for (int k = 0; k < 3; ++k) A.par_c[4 * r + 1 + k] = B[r][k] - A.N[k] * SB * A.inv_EN;
    }
    const double r1 = A.S[1] / A.S[0], r2 = A.S[2] / A.S[0];
    const double a1j = A.U[1] - r1 * A.U[0], a1i = A.V[1] - r1 * A.V[0];
    const double a2j = A.U[2] - r2 * A.U[0], a2i = A.V[2] - r2 * A.V[0];
    A.par_a[0] = (float)a1j;
    A.par_a[1] = (float)a1i;
    A.par_a[2] = (float)a2j;
    A.par_a[3] = (float)a2i;
    const double det = a1j * a2i - a1i * a2j;
    if (det == 0.0 || (float)a1j * (float)a2i - (float)a1i * (float)a2j == 0.0f) return;
    A.par_b[0] = (float)((std::fabs(a2i) + std::fabs(a1i)) / std::fabs(det));
    A.par_b[1] = (float)((std::fabs(a2j) + std::fabs(a1j)) / std::fabs(det));
}

// Dominant axis: argmax |e_c|, ties resolved in the order x, y, z.
int dominant_axis(V3 e)
{
    double ax = std::fabs(e[0]), ay = std::fabs(e[1]), az = std::fabs(e[2]);
    if (ax >= ay && ax >= az) return 0;
    if (ay >= az) return 1;
    return 2;
}

// FDK quantities of every angle (world coordinates; virtual detector through
// the volume centre c; angular steps of the source around the z axis through
// c from central differences of its unwrapped azimuth, one-sided at the ends).
void build_fdk(const tsx_geometry *g, tsx_plan *p)
{
    const int na = g->n_angles;
    V3 c;
    for (int k = 0; k < 3; ++k) c[k] = 0.5 * (g->window_min[k] + g->window_max[k]);
    std::vector<double> beta(na);
    for (int a = 0; a < na; ++a) {
        const double *vec = g->vectors + 12 * (int64_t)a;
        beta[a] = std::atan2(vec[1] - c[1], vec[0] - c[0]);
        if (a > 0) {  // unwrap
            while (beta[a] - beta[a - 1] > M_PI) beta[a] -= 2 * M_PI;
            while (beta[a] - beta[a - 1] < -M_PI) beta[a] += 2 * M_PI;
        }
    }
    p->fdk_host.resize(na);
    p->fdk_upright = true;
    for (int a = 0; a < na; ++a) {
        const double *vec = g->vectors + 12 * (int64_t)a;
        V3 s = mk(vec[0], vec[1], vec[2]), det = mk(vec[3], vec[4], vec[5]);
        V3 u = mk(vec[6], vec[7], vec[8]), v = mk(vec[9], vec[10], vec[11]);
        V3 n = cross(u, v);
        n = (1.0 / std::sqrt(dot(n, n))) * n;
        if (dot(det - s, n) < 0) n = -1.0 * n;
        V3 P0 = det + (0.5 - 0.5 * g->det_col_count) * u + (0.5 - 0.5 * g->det_row_count) * v;
        const double lam = dot(c - s, n) / dot(P0 - s, n);
        V3 P0s = lam * (P0 - s), U = lam * u, Vv = lam * v;
        V3 VxN = cross(Vv, n), UxN = cross(U, n);
        V3 Bj = (1.0 / dot(VxN, U)) * VxN, Bi = (1.0 / dot(UxN, Vv)) * UxN;
        const double vn = std::sqrt(dot(v, v));
        if (std::fabs(v[0]) > 1e-12 * vn || std::fabs(v[1]) > 1e-12 * vn || std::fabs(n[2]) > 1e-12)
            p->fdk_upright = false;
        FdkAngle &F = p->fdk_host[a];
        for (int k = 0; k < 3; ++k) {
            F.s[k] = s[k];
            F.n[k] = n[k];
            F.P0s[k] = P0s[k];
            F.U[k] = U[k];
            F.V[k] = Vv[k];
            F.Bj[k] = Bj[k];
            F.Bi[k] = Bi[k];
        }
        F.D = dot(c - s, n);
        F.cj = -dot(P0s, Bj);
        F.ci = -dot(P0s, Bi);
        double db;
        if (na < 2)
            db = 2 * M_PI;
        else if (a == 0)
            db = std::fabs(beta[1] - beta[0]);
        else if (a == na - 1)
            db = std::fabs(beta[na - 1] - beta[na - 2]);
        else
            db = 0.5 * std::fabs(beta[a + 1] - beta[a - 1]);
        F.half_dbeta = 0.5 * db;
        F.inv_du = 1.0 / std::sqrt(dot(U, U));
    }
}

// The device angle table of a plan: the angles in order, then ordered by
// stepping axis with their original index in `pad`.
std::vector<AngleParams> device_angle_table(const std::vector<AngleParams> &host)
{
    std::vector<AngleParams> t(host);
    for (int c = 0; c < 3; ++c)
        for (size_t a = 0; a < host.size(); ++a)
            if (host[a].m == c) {
                t.push_back(host[a]);
                t.back().pad = (int32_t)a;
            }
    return t;
}

int build_angles(const tsx_geometry *g, tsx_plan *p)
{
    const int64_t n3[3] = {g->grid_col_count, g->grid_row_count, g->grid_slice_count};
    const int64_t stride3[3] = {1, n3[0], n3[0] * n3[1]};
    double h[3];
    for (int c = 0; c < 3; ++c) h[c] = (g->window_max[c] - g->window_min[c]) / (double)n3[c];
    const V3 centre = mk(0.5 * (n3[0] - 1), 0.5 * (n3[1] - 1), 0.5 * (n3[2] - 1));
    const bool cone = g->type == TSX_CONE_VEC;
    p->host.resize(g->n_angles);
    std::vector<double> slope_j(g->n_angles), slope_i(g->n_angles);  // tile_slopes per angle
    for (int a = 0; a < g->n_angles; ++a) {
        const double *vec = g->vectors + 12 * (int64_t)a;
        V3 first, det, u, v;
        for (int c = 0; c < 3; ++c) {
            first[c] = vec[c];
            det[c] = vec[3 + c];
            u[c] = vec[6 + c];
            v[c] = vec[9 + c];
        }
        if (!finite3(first) || !finite3(det) || !finite3(u) || !finite3(v))
            return fail(TSX_ERR_INVALID_ARGUMENT,
                        "geometry vectors of angle " + std::to_string(a) + " are not finite");
        V3 U, Vv, Dg, S, e;
        for (int c = 0; c < 3; ++c) {
            U[c] = u[c] / h[c];
            Vv[c] = v[c] / h[c];
            Dg[c] = (det[c] - g->window_min[c]) / h[c] - 0.5;
            if (cone) {
                S[c] = (first[c] - g->window_min[c]) / h[c] - 0.5;
                e[c] = (det[c] - first[c]) / h[c];
            } else {
                S[c] = first[c] / h[c];
                e[c] = S[c];
            }
        }
        // centre of pixel (0, 0)
        V3 P0 = Dg + (0.5 - 0.5 * g->det_col_count) * U + (0.5 - 0.5 * g->det_row_count) * Vv;
        V3 N = cross(U, Vv);
        double nn = std::sqrt(dot(N, N));
        if (!(nn > 0.0))
            return fail(TSX_ERR_INVALID_ARGUMENT,
                        "detector u and v of angle " + std::to_string(a) + " are parallel or zero");
        double ee = std::sqrt(dot(e, e));
        if (!(ee > 0.0))
            return fail(TSX_ERR_INVALID_ARGUMENT,
                        cone ? "source coincides with detector centre at angle " + std::to_string(a)
                             : "zero ray direction at angle " + std::to_string(a));
        // move the detector to a parallel plane through the volume centre
        if (cone) {
            double den = dot(P0 - S, N);
            if (std::fabs(den) <= 1e-12 * nn * std::sqrt(dot(P0 - S, P0 - S)))
                return fail(TSX_ERR_INVALID_ARGUMENT,
                            "source lies in the detector plane at angle " + std::to_string(a));
            double lam = dot(centre - S, N) / den;
            if (std::fabs(lam) < 1e-6) lam = 1.0;  // source in the central plane: keep the real detector
            P0 = S + lam * (P0 - S);
            U = lam * U;
            Vv = lam * Vv;
        } else {
            double en = dot(S, N);
            if (std::fabs(en) <= 1e-12 * nn * ee)
                return fail(TSX_ERR_INVALID_ARGUMENT,
                            "ray direction is parallel to the detector at angle " + std::to_string(a));
            double t0 = dot(centre - P0, N) / en;
            P0 = P0 + t0 * S;
        }
        N = cross(U, Vv);
        V3 VxN = cross(Vv, N), UxN = cross(U, N);
        V3 Bj = (1.0 / dot(VxN, U)) * VxN;
        V3 Bi = (1.0 / dot(UxN, Vv)) * UxN;

        AngleParams &A = p->host[a];
        std::memset(&A, 0, sizeof(A));
        int m = dominant_axis(e);
        int c1 = (m == 2) ? 1 : 2;  // larger-stride in-slice axis
        int c2 = (m == 0) ? 1 : 0;  // smaller-stride in-slice axis
        // (m, c1, c2): x -> (x, z, y), y -> (y, z, x), z -> (z, y, x)
        const int ax[3] = {m, c1, c2};
        for (int k = 0; k < 3; ++k) {
            int c = ax[k];
            A.P0[k] = P0[c];
            A.U[k] = U[c];
            A.V[k] = Vv[c];
            A.S[k] = S[c];
            A.N[k] = N[c];
            A.Bj[k] = Bj[c];
            A.Bi[k] = Bi[c];
            A.h[k] = h[c];
            A.n[k] = (int32_t)n3[c];
            A.st[k] = stride3[c];
        }
        {
            const int64_t P2 = 2 * PAD;
            if (m == 0) {  // [nz+2P][nx][ny+2P]
                A.pst[0] = n3[1] + P2;
                A.pst[1] = n3[0] * (n3[1] + P2);
            } else if (m == 1) {  // [nz+2P][ny][nx+2P]
                A.pst[0] = n3[0] + P2;
                A.pst[1] = n3[1] * (n3[0] + P2);
            } else {  // [nz][ny+2P][nx+2P]
                A.pst[0] = (n3[1] + P2) * (n3[0] + P2);
                A.pst[1] = n3[0] + P2;
            }
            A.pst[2] = 1;
        }
        A.m = m;
        A.c1 = c1;
        A.c2 = c2;
        A.cone = cone ? 1 : 0;
        A.kappa = cone ? dot(P0 - S, N) : 0.0;
        A.EN = cone ? 0.0 : dot(S, N);
        A.inv_kappa = A.kappa != 0.0 ? 1.0 / A.kappa : 0.0;
        A.inv_EN = A.EN != 0.0 ? 1.0 / A.EN : 0.0;
        A.inv_S0 = !cone && A.S[0] != 0.0 ? 1.0 / A.S[0] : 0.0;
        for (int k = 0; k < 3; ++k) {
            A.fU[k] = (float)A.U[k];
            A.fV[k] = (float)A.V[k];
        }
        parallel_map(A);
        tile_slopes(A, g->det_row_count, g->det_col_count, &slope_j[a], &slope_i[a]);
        p->axis_count[m] += 1;
    }
    // forward tiles: waves along detector columns where every angle's rays
    // keep to a slice row along a direction closer to v than to u, else along
    // rows; the largest offset across a wave, dropped where it would reach
    // half the detector's extent across the waves
    p->fp_column_waves = p->n_angles > 0;
    for (int a = 0; a < p->n_angles; ++a)
        p->fp_column_waves = p->fp_column_waves && std::fabs(slope_j[a]) > std::fabs(slope_i[a]);
    p->fp_shear_rows = 0;
    for (int a = 0; a < p->n_angles; ++a) {
        AngleParams &A = p->host[a];
        A.shear = tile_shear(slope_j[a], slope_i[a], p->fp_column_waves);
        p->fp_shear_rows =
            std::max(p->fp_shear_rows, (int32_t)std::ceil(32.0 * std::fabs((double)A.shear) - 1e-9));
    }
    if (2 * p->fp_shear_rows > (p->fp_column_waves ? p->cols : p->rows)) p->fp_shear_rows = 0;
    for (int c = 0; c < 3; ++c) {
        p->vlo[c] = g->window_min[c];
        p->vh[c] = h[c];
    }
    if (cone) build_fdk(g, p);
    // padded layouts of the forward projector
    const int64_t P2 = 2 * PAD;
    p->layout_elems[0] = (n3[2] + P2) * n3[0] * (n3[1] + P2);
    p->layout_elems[1] = (n3[2] + P2) * n3[1] * (n3[0] + P2);
    p->layout_elems[2] = n3[2] * (n3[1] + P2) * (n3[0] + P2);
    // elements between rows of c1 in each layout; the march forms row offsets
    // with 24-bit multiplies, or with 32-bit ones where a row-paired row
    // reaches 2^24 bytes (fp_wide: slices of 2048 x 2048 and up)
    const int64_t row1[3] = {n3[0] * (n3[1] + P2), n3[1] * (n3[0] + P2), n3[0] + P2};
    p->fp_padded = true;
    p->fp_pair = true;
    p->fp_wide = false;
    for (int c = 0; c < 3; ++c)
        if (p->axis_count[c] > 0 && row1[c] * 8 >= ((int64_t)1 << 24)) p->fp_wide = true;
    // The march addresses a layout with 32-bit byte offsets. A volume whose
    // layouts reach chunk_bytes is projected in z-slabs of at most `depth`
    // slices: the layouts of a slab of depth d hold (d + 2 PAD) K_c elements
    // (x, y stepping) or d K_z (z stepping). The depth is chosen for row-paired
    // elements where those fit, so that every march variant can take the slabs.
    p->fp_slab_depth = p->nz;
    if (p->fp_padded && !p->is_slab) {
        const int64_t K[3] = {n3[0] * (n3[1] + P2), n3[1] * (n3[0] + P2), (n3[1] + P2) * (n3[0] + P2)};
        auto deepest = [&](int es) {
            int64_t d = n3[2];
            for (int c = 0; c < 3; ++c)
                if (p->axis_count[c] > 0)
                    d = std::min(d, (p->chunk_bytes - 1) / (4 * es * K[c]) - (c < 2 ? P2 : 0));
            return d;
        };
        int64_t d = deepest(p->fp_pair ? 2 : 1);
        if (d < 1 && p->fp_pair) {
            p->fp_pair = false;
            d = deepest(1);
        }
        if (d < 1) {
            p->fp_padded = false;
        } else if (d < n3[2]) {
            const int64_t n_slabs = (n3[2] + d - 1) / d;
            p->fp_slab_depth = (int32_t)((n3[2] + n_slabs - 1) / n_slabs);
        }
    }
    // separable backprojection: upright detector (v along z), stepping along x or y
    p->bp_separable = g->det_supersampling == 1;
    if (!p->bp_separable) p->bp_reason = "detector supersampling";
    if (g->det_col_count > BP_RT_MAX_COLS) {
        p->bp_separable = false;
        p->bp_reason = "more than " + std::to_string(BP_RT_MAX_COLS) + " detector columns";
    }
    // transposed projections of one angle (the separable backprojector's
    // workspace); a launch takes as many angles as stay below chunk_bytes
    {
        const int64_t per_angle = (int64_t)(g->det_row_count + 2 * RPAD) * g->det_col_count * 4;
        p->bp_chunk = (int32_t)std::min<int64_t>(g->n_angles, (p->chunk_bytes - 1) / per_angle);
        if (p->bp_chunk < 1) {
            p->bp_separable = false;
            p->bp_reason = "the projections of one angle reach " + std::to_string(p->chunk_bytes) + " bytes";
        }
    }
    for (int a = 0; a < g->n_angles && p->bp_separable; ++a) {
        const AngleParams &A = p->host[a];
        const double vn = std::sqrt(A.V[0] * A.V[0] + A.V[1] * A.V[1] + A.V[2] * A.V[2]);
        if (A.m == 2) {
            p->bp_separable = false;
            p->bp_reason = "angle " + std::to_string(a) + " steps along z";
        } else if (std::fabs(A.V[0]) > 1e-12 * vn || std::fabs(A.V[2]) > 1e-12 * vn) {
            p->bp_separable = false;
            p->bp_reason = "detector v is not parallel to z at angle " + std::to_string(a);
        } else if (A.cone) {
            // the source must lie outside the slab of slices (with margin), and every
            // voxel must project with a positive scale onto the detector plane
            // bricks cover the volume rounded up to whole bricks
            const double nm = (double)((A.n[0] + BR - 1) / BR * BR);
            const double nc = (double)((A.n[2] + BR - 1) / BR * BR);
            if (!(A.S[0] < -2.0 || A.S[0] > nm + 1.0)) {
                p->bp_separable = false;
                p->bp_reason = "source within the slice range at angle " + std::to_string(a);
            }
            for (int cm = 0; cm < 2 && p->bp_separable; ++cm)
                for (int cc = 0; cc < 2 && p->bp_separable; ++cc) {
                    const double X0 = cm ? nm + 1.0 : -2.0, X2 = cc ? nc + 1.0 : -2.0;
                    const double dn = (X0 - A.S[0]) * A.N[0] + (X2 - A.S[2]) * A.N[2];
                    if (!(A.kappa / dn > 0.0)) {
                        p->bp_separable = false;
                        p->bp_reason = "volume not in front of the source at angle " + std::to_string(a);
                    }
                }
        }
    }
    return TSX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------
namespace {

int check_plan(const tsx_plan *p)
{
    if (!p) return fail(TSX_ERR_INVALID_ARGUMENT, "plan is NULL");
    if (p->device < 0 || !p->dev)
        return fail(TSX_ERR_NO_DEVICE, "plan was created without a device (device = -1)");
    return TSX_OK;
}

struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(dev) == hipSuccess) ok = true;
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Integer value of an environment variable, or -1 when it is not set.
int env_int(const char *name)
{
    const char *e = std::getenv(name);
    return (e && e[0]) ? std::atoi(e) : -1;
}

// Kernel variant of the padded forward projector: 2 = each lane marches its
// own ray's slice range, four samples in flight; 3 = the lanes of a wave march
// the same slices (coalesced reads); 4 = as 3 on row-paired layouts (one
// 16-byte load per sample; default; falls back to 3 where those layouts would
// not fit 32-bit offsets). TSX_FP_VARIANT overrides the plan's choice.
int fp_variant_of(const tsx_plan *p)
{
    const int e = env_int("TSX_FP_VARIANT");
    int v = (e >= 2 && e <= 5) ? e : (p->fp_variant >= 0 ? p->fp_variant : 4);
    if (p->fp_wide && v != 3) v = 4;  // (wide rows: marches 3 and 4 only)
    if (v == 4 && !p->fp_pair) v = 3;
    return v;
}

// Floats per element of the forward projector's padded layouts.
int fp_esize(const tsx_plan *p) { return fp_variant_of(p) == 4 ? 2 : 1; }

// Detector rows per block of the padded forward projector: 8 (default), or
// 4 or 16 for marches 3 and 4 without supersampling. TSX_FP_TILE_ROWS
// overrides the plan's choice.
int fp_tile_rows_of(const tsx_plan *p)
{
    const int v = fp_variant_of(p);
    if (p->ss != 1 || p->fp_wide || (v != 3 && v != 4)) return 8;
    const int e = env_int("TSX_FP_TILE_ROWS");
    if (e == 4 || e == 8 || e == 16) return e;
    return p->fp_by > 0 ? p->fp_by : 8;
}

// Row offset bound of the forward projector's sheared tiles as launched: the
// plan's (0 without shear) unless the plan's choice or TSX_FP_SHEAR=0 turns
// them off (TSX_FP_SHEAR=1 turns them on over a tuned choice).
bool fp_shear_on(const tsx_plan *p)
{
    const int e = env_int("TSX_FP_SHEAR");
    return e == 0 ? false : (e == 1 ? true : p->fp_shear != 0);
}

int fp_shear_rows_of(const tsx_plan *p) { return fp_shear_on(p) ? p->fp_shear_rows : 0; }

// Largest byte size of one padded forward layout and of the transposed
// projections of one backprojection launch: the kernels address them with
// 32-bit byte offsets. TSX_CHUNK_BYTES, read when a plan is created, lowers
// it (tests reach the slab and angle-chunk paths at small sizes with it).
int64_t chunk_limit()
{
    const int64_t full = (int64_t)1 << 32;
    const char *e = std::getenv("TSX_CHUNK_BYTES");
    if (e && e[0]) {
        const long long v = std::atoll(e);
        if (v > 0 && v < full) return v;
    }
    return full;
}

// The epilogue of the second and later chunks (slabs of the volume, ranges
// of angles) of a launch whose first chunk takes `e`: the chunks' sums add
// up in the output.
Epilogue later_chunk(Epilogue e)
{
    if (e.mode == EP_SET) {
        e.mode = EP_ADD;
    } else if (e.mode == EP_RESIDUAL) {
        e.mode = EP_SCALED_SUB;
        e.y = nullptr;
    }
    return e;
}

// z-extent of the separable backprojector's bricks: 32 (default), 24 or 16.
// TSX_BP_BRICK_Z overrides the plan's choice.
int bp_brick_z_of(const tsx_plan *p)
{
    const int e = env_int("TSX_BP_BRICK_Z");
    if (e == 16 || e == 24 || e == 32) return e;
    return p->bp_bz > 0 ? p->bp_bz : 32;
}

bool force_general()
{
    const char *e = std::getenv("TSX_FORCE_GENERAL");
    return e && e[0] == '1';
}

// Stream-ordered scratch memory, returned to the device pool when the work
// on `stream` that uses it has finished.
struct Scratch {
    void *ptr = nullptr;
    hipStream_t stream = nullptr;
    int alloc(size_t bytes, hipStream_t s)
    {
        stream = s;
        hipError_t e = hipMallocAsync(&ptr, bytes, s);
        if (e != hipSuccess) {
            ptr = nullptr;
            return fail(TSX_ERR_OUT_OF_MEMORY, std::string("hipMallocAsync of ") +
                                                   std::to_string(bytes) + " bytes failed: " +
                                                   hipGetErrorString(e));
        }
        return TSX_OK;
    }
    ~Scratch()
    {
        if (ptr) (void)hipFreeAsync(ptr, stream);
    }
};

void keep_pool_memory(int device)
{
    static bool done[64] = {false};
    if (device < 0 || device >= 64 || done[device]) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    done[device] = true;
}

// Floats per element of the padded layouts of forward variant `variant`.
int esize_of(int variant) { return variant == 4 ? 2 : 1; }

// Scratch floats of one batch entry of the fast forward projector: the
// padded layouts of the volume (of its largest slab when it is projected in
// slabs, the slabs taking turns in the same memory).
int64_t fp_workspace_elems(const tsx_plan *p)
{
    const int es = fp_esize(p);
    int64_t per_entry = 0;
    for (const tsx_plan *q : p->slabs.empty() ? std::vector<const tsx_plan *>{p}
                                              : std::vector<const tsx_plan *>(p->slabs.begin(), p->slabs.end())) {
        int64_t e = 0;
        for (int c = 0; c < 3; ++c)
            if (q->axis_count[c] > 0) e += q->layout_elems[c] * es;
        per_entry = std::max(per_entry, e);
    }
    return per_entry;
}

// Sorted angles per launch of the separable backprojector.
int bp_chunk_of(const tsx_plan *p) { return std::max(1, std::min(p->bp_chunk, p->n_angles)); }

int bp_brick_z_of(const tsx_plan *p);

// Angle parts of the separable backprojector. A block owns a brick and loops
// over the angles, so a volume of few bricks (64^3: 32, 128^3: 256) leaves
// most CUs idle or holds one block per CU; its angles are then split into
// parts of part_len >= 4 consecutive sorted angles, so that bricks x parts
// reaches about 1024 blocks (four per CU). Single launches run
// the parts in separate blocks and add their sums in order (bp_parts_reduce);
// fused launches run them in turn in one block and add them in the same order,
// so both give the same sums. Volumes of 512 bricks or more, and plans whose
// angles take several launches, keep one part. TSX_BP_PARTS=1 keeps one part
// everywhere.
int bp_part_len_of(const tsx_plan *p)
{
    if (!p->bp_separable || bp_chunk_of(p) < p->n_angles || env_int("TSX_BP_PARTS") == 1) return p->n_angles;
    const int bz = bp_brick_z_of(p);
    const int64_t bricks = (int64_t)((p->nx + BR - 1) / BR) * ((p->ny + BR - 1) / BR) * ((p->nz + bz - 1) / bz);
    if (bricks >= 512) return p->n_angles;
    const int64_t len = std::max<int64_t>(4, ((int64_t)p->n_angles * bricks + 1023) / 1024);
    return (int)std::min<int64_t>(len, p->n_angles);
}
int bp_parts_of(const tsx_plan *p)
{
    const int len = bp_part_len_of(p);
    return (p->n_angles + len - 1) / len;
}

// Length-weighted projections of one batch entry: transposed and padded
// ([sorted angle][col][RPAD + row], scale_by_length_t) for the separable
// kernel, for the angles of one launch; in the projections' own layout for
// the general one.
int64_t bp_workspace_elems(const tsx_plan *p)
{
    if (p->bp_separable) return (int64_t)bp_chunk_of(p) * p->cols * (p->rows + 2 * RPAD);
    return (int64_t)p->rows * p->n_angles * p->cols;
}

// The general backprojector takes length-weighted projections (like the
// separable one) for detector supersampling 1, within the launch limits of
// the weighting pass.
bool bp_general_prescaled(const tsx_plan *p, int64_t /*batch*/)
{
    return !p->bp_separable && p->ss == 1;
}

// Scratch sizes rounded up to 256 bytes.
int64_t align_ws(int64_t v) { return (v + 255) / 256 * 256; }

// The forward projector of a plan with an axis-swapped plan runs on that
// plan (the volume moved into its layout first) when that plan has the padded
// kernels: its detector u axis lies along the layouts' rows, where a wave's
// rays read consecutive elements.
bool fp_swapped(const tsx_plan *p) { return p->perm && p->perm->fp_padded && !force_general(); }

// Projections of a plan with detector supersampling run on its
// refined-detector plan where that plan has the fast kernels: the
// backprojection where it has the separable one, the forward projection
// where it has the padded one (directly or through its axis swap).
bool bp_refined(const tsx_plan *p)
{
    return p->fine && (p->fine->bp_separable || p->fine->perm) && !force_general();
}
// Elements of one batch entry's projections on the refined detector.
int64_t refined_sino(const tsx_plan *p)
{
    return (int64_t)p->fine->rows * p->fine->n_angles * p->fine->cols;
}

// (the refined forward path holds the refined projections in its workspace:
// used while they stay within 16 GiB per batch entry, the native supersampled
// kernels beyond)
constexpr int64_t REFINED_FP_MAX_BYTES = (int64_t)16 << 30;

bool fp_refined(const tsx_plan *p)
{
    return p->fine && (p->fine->fp_padded || fp_swapped(p->fine)) && !force_general() &&
           refined_sino(p) * 4 <= REFINED_FP_MAX_BYTES;
}

int64_t workspace_bytes(const tsx_plan *p, int op, int64_t batch)
{
    if (force_general()) return 0;
    // the axis-swapped plan's volume, then that plan's workspace
    if (op == TSX_OP_FP && fp_swapped(p))
        return align_ws((int64_t)p->nx * p->ny * p->nz * batch * 4) + workspace_bytes(p->perm, op, batch);
    // the refined detector's projections, then that plan's workspace
    if (op == TSX_OP_FP && fp_refined(p))
        return align_ws(refined_sino(p) * batch * 4) + workspace_bytes(p->fine, op, batch);
    if (op == TSX_OP_FP) return p->fp_padded ? fp_workspace_elems(p) * batch * 4 : 0;
    if (op == TSX_OP_BP && bp_refined(p)) return workspace_bytes(p->fine, op, batch);
    if (p->perm)
        return align_ws((int64_t)p->nx * p->ny * p->nz * batch * 4) + workspace_bytes(p->perm, op, batch);
    if (p->bp_separable) {
        // transposed projections, then the sums of the angle parts
        const int parts = bp_parts_of(p);
        const int64_t t = bp_workspace_elems(p) * batch * 4;
        return parts > 1 ? align_ws(t) + (int64_t)parts * p->nx * p->ny * p->nz * batch * 4 : t;
    }
    return bp_general_prescaled(p, batch) ? bp_workspace_elems(p) * batch * 4 : 0;
}

// Launches of one block per logical id take the chunked XCD order of
// xcd_logical: chunks of 2^shift consecutive ids, the largest power of two up
// to 1024 that still gives every XCD about 256 chunks, and a grid with the same
// number of chunk slots on every XCD (blocks past the last id return at once).
// Single 512^3 instance, forward: 45.4 ms with ~64 chunks per XCD, 41.5 ms
// with ~256 (the chunks' costs differ, more of them even out the XCDs).
// Logical ids stay below MAX_LOGICAL so that the grid fits in an int, and the
// grid's work-items (blocks x threads per block) below 2^32, the limit of the
// dispatch packet: max_blocks() gives the bound for a block size (the grid
// rounds the ids up by less than 9 x 1024).
constexpr int64_t MAX_LOGICAL = 0x7fffffff - 9 * 1024;

int64_t max_blocks(int threads_per_block)
{
    return std::min(MAX_LOGICAL, (int64_t)(0xffffffffLL / threads_per_block) - 9 * 1024);
}

unsigned xcd_grid(int64_t n, int *shift)
{
    int s = 0;
    while (s < 10 && ((int64_t)2048 << (s + 1)) <= n) ++s;
    const int64_t c = (int64_t)1 << s;
    const int64_t chunks = (n + c - 1) / c;
    *shift = s;
    return (unsigned)((chunks + 7) / 8 * 8 * c);
}

// The epilogue of batch entries b0, b0 + 1, ... of a launch that starts at
// entry b0 of the caller's batch.
Epilogue ep_from(Epilogue e, int64_t b0)
{
    if (e.y) e.y += b0 * e.stride;
    if (e.w) e.w += b0 * e.stride;
    return e;
}

// Grid limits of the launches with one grid dimension per batch entry, angle
// or slice (65535 per dimension).
constexpr int64_t GRID_DIM_MAX = 65535;

// The length-weighted projections ly = l * y of `batch` entries.
void launch_scale_by_length(const tsx_plan *p, const float *proj, int64_t ps, float *ly, int64_t ly_stride,
                            int64_t batch, hipStream_t stream)
{
    const int64_t lines = (int64_t)p->rows * p->n_angles * batch;
    dim3 grid((unsigned)((p->cols + 255) / 256), (unsigned)std::min(lines, GRID_DIM_MAX));
    hipLaunchKernelGGL(scale_by_length, grid, dim3(256), 0, stream, p->dev, p->n_angles, p->rows, p->cols,
                       proj, ps, ly, ly_stride, lines);
}

// Allocates and fills the slice-range table of a separable plan whose device
// angle table is in place (work enqueued on `stream`).
int build_range_table(tsx_plan *p, hipStream_t stream, uint32_t **out)
{
    *out = nullptr;
    if (!p->bp_separable) return TSX_OK;
    const int64_t stride = bp_rt_stride(p->nx, p->ny);
    const hipError_t e = hipMalloc((void **)out, (size_t)(stride * p->n_angles * 4));
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(TSX_ERR_OUT_OF_MEMORY, std::string("slice-range table: ") + hipGetErrorString(e));
    }
    p->bp_rt_stride = stride;
    dim3 grid((unsigned)((stride + 255) / 256), (unsigned)p->n_angles);
    hipLaunchKernelGGL(bp_range_table, grid, dim3(256), 0, stream, p->dev + p->n_angles, p->cols, p->nx, p->ny,
                       stride, *out);
    TSX_HIP_CHECK(hipGetLastError());
    return TSX_OK;
}

// The transposed, padded length-weighted projections of the sorted angles
// k0 .. k0 + kn - 1 of `batch` entries.
void launch_scale_by_length_t(const tsx_plan *p, int k0, int kn, const float *proj, int64_t ps, float *ly,
                              int64_t ly_stride, int64_t batch, hipStream_t stream, int up = 1)
{
    const int64_t lines = (int64_t)kn * batch;
    const unsigned tiles = (unsigned)(((p->rows + 31) / 32) * ((p->cols + 31) / 32));
    dim3 grid(tiles, (unsigned)std::min(lines, GRID_DIM_MAX));
    hipLaunchKernelGGL(scale_by_length_t, grid, dim3(32, 8), 0, stream, p->dev + p->n_angles + k0, kn,
                       p->n_angles, p->rows, p->cols, proj, ps, ly, ly_stride, lines, up);
}

// Either the caller's workspace (checked for size) or stream-ordered memory.
int get_workspace(const tsx_plan *p, int64_t bytes, void *ws, int64_t ws_bytes, hipStream_t stream,
                  Scratch &owned, void **out)
{
    if (ws) {
        if (ws_bytes < bytes)
            return fail(TSX_ERR_INVALID_ARGUMENT, "workspace of " + std::to_string(ws_bytes) +
                                                      " bytes is smaller than the " +
                                                      std::to_string(bytes) + " bytes required");
        *out = ws;
        return TSX_OK;
    }
    keep_pool_memory(p->device);
    int rc = owned.alloc((size_t)bytes, stream);
    *out = owned.ptr;
    return rc;
}

// Offsets (in floats) of the layouts of one volume, and their total.
void layout_offsets(const tsx_plan *p, int es, int64_t offs[3], int64_t *per_entry)
{
    int64_t e = 0;
    for (int c = 0; c < 3; ++c) {
        offs[c] = e;
        if (p->axis_count[c] > 0) e += p->layout_elems[c] * es;
    }
    *per_entry = e;
}

// Entries of a job table past job n - 1 (and its end begin[n]): INT_MAX, as
// owner_of() expects.
template <int N>
void pad_begin(int (&begin)[N], int n)
{
    for (int e = n + 1; e < N; ++e) begin[e] = 0x7fffffff;
}

// Padded layouts of several volumes, built by build_layouts_multi in as few
// launches as its job lists and the dispatch limit allow: one list of x
// layouts (transposes), one of y and z layouts (copies). add() collects,
// flush() launches what was collected.
struct LayoutBatch {
    LayoutJobs L[2];  // [0]: copies, [1]: transposes
    int maxb[2] = {0, 0};
    hipStream_t stream;
    explicit LayoutBatch(hipStream_t s) : stream(s)
    {
        for (LayoutJobs &l : L) l.n = 0;
    }
    void add(const tsx_plan *p, int es, const float *src, float *dst, const int64_t offs[3])
    {
        for (int c = 0; c < 3; ++c) {
            if (p->axis_count[c] == 0) continue;
            LayoutJob j;
            j.src = src;
            j.dst = dst + offs[c];
            j.nx = p->nx;
            j.ny = p->ny;
            j.nz = p->nz;
            j.axis = c;
            j.pair = es == 2;
            j.blocks = (int)layout_job_blocks(j);
            const int t = c == 0;
            LayoutJobs &l = L[t];
            // a grid of (largest job) x (jobs) blocks of 256 stays below the
            // dispatch limit
            if (l.n == MAX_LAYOUT_JOBS ||
                (l.n > 0 && (int64_t)std::max(maxb[t], j.blocks) * (l.n + 1) > max_blocks(256)))
                launch(t);
            l.job[l.n] = j;
            maxb[t] = std::max(maxb[t], j.blocks);
            l.n += 1;
        }
    }
    void launch(int t)
    {
        LayoutJobs &l = L[t];
        if (l.n == 0) return;
        const dim3 grid((unsigned)maxb[t], (unsigned)l.n);
        if (t)
            hipLaunchKernelGGL(build_layouts_multi<true>, grid, dim3(32, 8), 0, stream, l);
        else
            hipLaunchKernelGGL(build_layouts_multi<false>, grid, dim3(32, 8), 0, stream, l);
        l.n = 0;
        maxb[t] = 0;
    }
    void flush()
    {
        launch(1);
        launch(0);
    }
};

// fp_general over every angle and batch entry, in launches of at most 65535
// angles and 65535 entries.
template <typename T>
int launch_fp_general(const tsx_plan *p, const T *vol, int64_t vs, T *proj, int64_t ps, int64_t batch,
                      Epilogue ep, hipStream_t stream)
{
    const int tiles_u = (p->cols + FP_BX - 1) / FP_BX;
    const int tiles_v = (p->rows + FP_BY - 1) / FP_BY;
    for (int64_t b0 = 0; b0 < batch; b0 += GRID_DIM_MAX) {
        const int64_t nb = std::min(GRID_DIM_MAX, batch - b0);
        for (int a0 = 0; a0 < p->n_angles; a0 += (int)GRID_DIM_MAX) {
            const int na = (int)std::min<int64_t>(GRID_DIM_MAX, p->n_angles - a0);
            dim3 grid((unsigned)(tiles_u * tiles_v), (unsigned)na, (unsigned)nb);
            hipLaunchKernelGGL(fp_general<T>, grid, dim3(FP_BX, FP_BY), 0, stream, p->dev, p->n_angles, a0,
                               p->rows, p->cols, p->ss, tiles_u, vol + b0 * vs, vs, proj + b0 * ps, ps,
                               ep_from(ep, b0));
        }
    }
    TSX_HIP_CHECK(hipGetLastError());
    return TSX_OK;
}

// The padded forward kernels by supersampling, march variant and tile rows
// (tiles of 4 and 16 rows for variants 3 and 4 without supersampling, 16
// only through TSX_FP_TILE_ROWS; 8 rows
// otherwise).
using FpKernel = void (*)(FpView, int, int);
using FpMultiKernel = void (*)(FpMulti);

template <int SS>
FpKernel fp_kernel_ss(int variant)
{
    if (variant == 4) return fp_padded<SS, 4, 8>;
    if (variant == 5) return fp_padded<SS, 5, 8>;
    if (variant == 3) return fp_padded<SS, 3, 8>;
    return fp_padded<SS, 2, 8>;
}

// Kernels for layouts whose rows reach 2^24 bytes: the wave-synchronous
// marches (row-paired or plain) on tiles of 8 rows.
template <int SS>
FpKernel fp_kernel_wide_ss(int variant)
{
    return variant == 4 ? fp_padded<SS, 4, 8, true> : fp_padded<SS, 3, 8, true>;
}

FpKernel fp_kernel(int ss, int variant, int by, bool wide = false)
{
    if (wide) {
        switch (ss) {
        case 1: return fp_kernel_wide_ss<1>(variant);
        case 2: return fp_kernel_wide_ss<2>(variant);
        case 3: return fp_kernel_wide_ss<3>(variant);
        default: return fp_kernel_wide_ss<4>(variant);
        }
    }
    if (ss == 1 && by == 4) return variant == 4 ? fp_padded<1, 4, 4> : fp_padded<1, 3, 4>;
    if (ss == 1 && by == 16) return variant == 4 ? fp_padded<1, 4, 16> : fp_padded<1, 3, 16>;
    switch (ss) {
    case 1: return fp_kernel_ss<1>(variant);
    case 2: return fp_kernel_ss<2>(variant);
    case 3: return fp_kernel_ss<3>(variant);
    default: return fp_kernel_ss<4>(variant);
    }
}

FpMultiKernel fp_multi_kernel(int variant, int by)
{
    if (by == 4) return variant == 4 ? fp_padded_multi<4, 4> : fp_padded_multi<3, 4>;
    if (by == 16) return variant == 4 ? fp_padded_multi<4, 16> : fp_padded_multi<3, 16>;
    if (variant == 4) return fp_padded_multi<4, 8>;
    if (variant == 5) return fp_padded_multi<5, 8>;
    if (variant == 3) return fp_padded_multi<3, 8>;
    return fp_padded_multi<2, 8>;
}

// The padded forward projector of variant `variant` with tiles of `by`
// detector rows on the whole volume of plan `p` (a plan without slabs, or
// one slab), its layouts built in `base`.
// shear: whether the tiles are sheared (p's own row offset bound then).
int launch_fp_padded(const tsx_plan *p, int variant, int by, bool shear, const float *vol, int64_t vs, float *proj,
                     int64_t ps, int64_t batch, Epilogue ep, float *base, hipStream_t stream)
{
    // tiles: FP_BX lanes along the wave (columns, or rows for column waves)
    // by BY waves
    const int along = p->fp_column_waves ? p->rows : p->cols, across = p->fp_column_waves ? p->cols : p->rows;
    const int tiles_u = (along + FP_BX - 1) / FP_BX;
    const int tiles_v = (across + by - 1) / by;
    const int es = esize_of(variant);
    // padded copies of every batch entry for each stepping axis in use
    int64_t per_entry = 0;
    int64_t offs[3] = {0, 0, 0};
    layout_offsets(p, es, offs, &per_entry);
    {
        LayoutBatch lb(stream);
        for (int64_t b = 0; b < batch; ++b) lb.add(p, es, vol + b * vs, base + b * per_entry, offs);
        lb.flush();
    }
    TSX_HIP_CHECK(hipGetLastError());
    // launches of whole batch entries while they fit the grid limit, else of
    // angle ranges of one entry
    const int64_t tiles = (int64_t)tiles_u * tiles_v;
    const int64_t cap = max_blocks(FP_BX * by);
    const int64_t per_b = tiles * p->n_angles;
    const int64_t bc = std::max<int64_t>(1, cap / per_b);      // entries per launch
    const int na_max = (int)std::min<int64_t>(p->n_angles, std::max<int64_t>(1, cap / tiles));
    for (int64_t b0 = 0; b0 < batch; b0 += bc) {
        const int64_t nbat = std::min(bc, batch - b0);
        for (int a0 = 0; a0 < p->n_angles; a0 += na_max) {
            FpView v;
            v.angles = p->dev;
            v.n_angles = p->n_angles;
            v.a0 = a0;
            v.na = std::min(na_max, p->n_angles - a0);
            v.rows = p->rows;
            v.cols = p->cols;
            v.tiles_u = tiles_u;
            v.tiles_v = tiles_v;
            v.shear_rows = shear ? p->fp_shear_rows : 0;
            v.column_waves = p->fp_column_waves ? 1 : 0;
            v.batch = (int)nbat;
            for (int c = 0; c < 3; ++c) {
                v.L.ptr[c] = p->axis_count[c] > 0 ? base + b0 * per_entry + offs[c] : nullptr;
                v.L.bytes[c] = (uint32_t)(p->layout_elems[c] * 4 * es);
            }
            v.L.stride = per_entry;
            v.proj = proj + b0 * ps;
            v.proj_stride = ps;
            v.ep = ep_from(ep, b0);
            const int n_blocks = (int)(tiles * v.na * nbat);
            int shift;
            const unsigned grid = xcd_grid(n_blocks, &shift);
            hipLaunchKernelGGL(fp_kernel(p->ss, variant, by, p->fp_wide), dim3(grid), dim3(FP_BX, by), 0, stream, v, n_blocks,
                               shift);
            TSX_HIP_CHECK(hipGetLastError());
        }
    }
    return TSX_OK;
}

Epilogue plain(int additive)
{
    Epilogue e;
    e.mode = additive ? EP_ADD : EP_SET;
    e.y = nullptr;
    e.w = nullptr;
    e.stride = 0;
    return e;
}

// out (extents nx, ny, nz; entries os floats apart) = src in the layout with
// axis swap - 1 and z exchanged (entries ss apart), through the epilogue; also
// the move of a volume into that layout (extents of the swapped layout).
int launch_unswap(const float *src, int64_t ss, float *out, int64_t os, int nx, int ny, int nz, int swap,
                  int64_t batch, Epilogue ep, hipStream_t stream)
{
    const int64_t blocks = swap == 1 ? (int64_t)((nx + 31) / 32) * ((nz + 31) / 32) * ny
                                     : (int64_t)((nx + 255) / 256) * ny * nz;
    const int64_t cap = max_blocks(256);
    for (int64_t b0 = 0; b0 < batch; b0 += GRID_DIM_MAX) {
        const int64_t nb = std::min(GRID_DIM_MAX, batch - b0);
        for (int64_t f = 0; f < blocks; f += cap)
            hipLaunchKernelGGL(unswap_volume, dim3((unsigned)std::min(cap, blocks - f), 1, (unsigned)nb),
                               dim3(32, 8), 0, stream, src + b0 * ss, ss, out + b0 * os, os, nx, ny, nz, swap, f,
                               ep_from(ep, b0));
    }
    TSX_HIP_CHECK(hipGetLastError());
    return TSX_OK;
}

// Forward projection of `batch` volumes: the padded kernels on the whole
// volume or slab by slab (each slab's sums added to the projections of the
// previous ones), else the general kernel.
int launch_fp(const tsx_plan *p, const float *vol, int64_t vs, float *proj, int64_t ps,
              int64_t batch, Epilogue ep, hipStream_t stream, void *wsp = nullptr,
              int64_t ws_bytes = 0)
{
    if (batch <= 0) return TSX_OK;
    if (fp_refined(p)) {
        // detector supersampling ss: the projections of the detector refined
        // ss times (its pixels are the sub-rays), averaged over each pixel's
        // ss x ss fine pixels into the caller's projections through the
        // epilogue
        // (the requirement formed here from its two parts, so that a short
        // workspace is refused rather than overrun)
        const int64_t fe = refined_sino(p);
        const int64_t fine_bytes = align_ws(fe * batch * 4), rest_bytes = workspace_bytes(p->fine, TSX_OP_FP, batch);
        Scratch owned;
        void *wptr = nullptr;
        int rc = get_workspace(p, fine_bytes + rest_bytes, wsp, ws_bytes, stream, owned, &wptr);
        if (rc) return rc;
        float *fine = (float *)wptr;
        rc = launch_fp(p->fine, vol, vs, fine, fe, batch, plain(0), stream, (char *)wptr + fine_bytes, rest_bytes);
        if (rc) return rc;
        const int64_t blocks = (int64_t)((p->cols + 255) / 256) * p->rows * p->n_angles;
        const int64_t cap = max_blocks(256);
        for (int64_t b0 = 0; b0 < batch; b0 += GRID_DIM_MAX) {
            const int64_t nb = std::min(GRID_DIM_MAX, batch - b0);
            for (int64_t f = 0; f < blocks; f += cap)
                hipLaunchKernelGGL(reduce_refined, dim3((unsigned)std::min(cap, blocks - f), 1, (unsigned)nb),
                                   dim3(256), 0, stream, fine + b0 * fe, fe, proj + b0 * ps, ps, p->rows,
                                   p->n_angles, p->cols, p->ss, f, ep_from(ep, b0));
        }
        TSX_HIP_CHECK(hipGetLastError());
        return TSX_OK;
    }
    if (fp_swapped(p)) {
        const int64_t ve = (int64_t)p->nx * p->ny * p->nz;
        const tsx_plan *q = p->perm;
        const int64_t tmp_bytes = align_ws(ve * batch * 4), rest_bytes = workspace_bytes(q, TSX_OP_FP, batch);
        Scratch owned;
        void *wptr = nullptr;
        int rc = get_workspace(p, tmp_bytes + rest_bytes, wsp, ws_bytes, stream, owned, &wptr);
        if (rc) return rc;
        float *tmp = (float *)wptr;
        rc = launch_unswap(vol, vs, tmp, ve, q->nx, q->ny, q->nz, p->swap, batch, plain(0), stream);
        if (rc) return rc;
        return launch_fp(q, tmp, ve, proj, ps, batch, ep, stream, (char *)wptr + tmp_bytes, rest_bytes);
    }
    if (!p->fp_padded || force_general() || p->ss > 4)
        return launch_fp_general<float>(p, vol, vs, proj, ps, batch, ep, stream);
    const int variant = fp_variant_of(p), by = fp_tile_rows_of(p);
    Scratch owned;
    void *wptr = nullptr;
    int rc = get_workspace(p, fp_workspace_elems(p) * batch * 4, wsp, ws_bytes, stream, owned, &wptr);
    if (rc) return rc;
    const bool shear = fp_shear_on(p);  // (slabs: each with its own row offset bound)
    if (p->slabs.empty())
        return launch_fp_padded(p, variant, by, shear, vol, vs, proj, ps, batch, ep, (float *)wptr, stream);
    for (size_t k = 0; k < p->slabs.size(); ++k) {
        rc = launch_fp_padded(p->slabs[k], variant, by, shear, vol + (int64_t)p->slab_z0[k] * p->ny * p->nx, vs,
                              proj, ps, batch, k == 0 ? ep : later_chunk(ep), (float *)wptr, stream);
        if (rc) return rc;
    }
    return TSX_OK;
}

// The separable backprojector over the sorted angles k0 .. k0 + kn - 1, their
// transposed projections in `ly` (batch entries ly_bytes apart).
BpView bp_view(const tsx_plan *p, int BZ, float *vol, int64_t vs, const float *ly, uint32_t ly_bytes,
               int batch, Epilogue ep, int k0, int kn)
{
    BpView v;
    v.angles = p->dev;
    v.sorted = p->dev + p->n_angles + k0;
    v.rt = p->bp_rt + (int64_t)k0 * p->bp_rt_stride;
    v.rt_stride = p->bp_rt_stride;
    v.n_angles = kn;
    v.rows = p->rows;
    v.cols = p->cols;
    v.nx = p->nx;
    v.ny = p->ny;
    v.nz = p->nz;
    v.bricks_x = (p->nx + BR - 1) / BR;
    v.bricks_y = (p->ny + BR - 1) / BR;
    v.bricks_z = (p->nz + BZ - 1) / BZ;
    v.batch = batch;
    v.vol = vol;
    v.vol_stride = vs;
    v.ly = ly;
    v.ly_bytes = ly_bytes;
    v.ep = ep;
    v.parts = 1;
    v.part_len = kn;
    v.partial = nullptr;
    return v;
}

// bp_general over every batch entry, in launches of at most 65535 slices x
// entries (grid z); `ly` non-null: length-weighted projections (PRESCALED).
template <typename T, bool PRESCALED>
int launch_bp_general(const tsx_plan *p, T *vol, int64_t vs, const T *proj, int64_t ps, int64_t batch,
                      Epilogue ep, hipStream_t stream)
{
    if (p->nz > GRID_DIM_MAX) return fail(TSX_ERR_UNSUPPORTED, "volume depth exceeds 65535 slices");
    const int64_t bc = GRID_DIM_MAX / p->nz;  // entries per launch
    for (int64_t b0 = 0; b0 < batch; b0 += bc) {
        const int64_t nb = std::min(bc, batch - b0);
        dim3 grid((unsigned)((p->nx + BP_BX - 1) / BP_BX), (unsigned)((p->ny + BP_BY - 1) / BP_BY),
                  (unsigned)(p->nz * nb));
        auto kernel = p->kind == TSX_CONE_VEC ? bp_general<T, PRESCALED, true> : bp_general<T, PRESCALED, false>;
        hipLaunchKernelGGL(kernel, grid, dim3(BP_BX, BP_BY), 0, stream, p->dev,
                           p->n_angles, p->rows, p->cols, p->ss, p->nx, p->ny, p->nz, vol + b0 * vs, vs,
                           proj + b0 * ps, ps, ep_from(ep, b0));
    }
    TSX_HIP_CHECK(hipGetLastError());
    return TSX_OK;
}


// up > 1: p is the refined-detector plan of a plan with detector
// supersampling up, and proj holds that plan's (coarse) projections.
int launch_bp(const tsx_plan *p, float *vol, int64_t vs, const float *proj, int64_t ps,
              int64_t batch, Epilogue ep, hipStream_t stream, void *wsp = nullptr,
              int64_t ws_bytes = 0, int up = 1)
{
    if (batch <= 0) return TSX_OK;
    if (bp_refined(p)) {
        // detector supersampling ss: the backprojection of the ss x ss
        // sub-rays of every pixel is that of the detector refined ss times
        // (the sub-rays are its pixels' rays), each fine pixel taking its
        // coarse pixel's value over ss^2 (in the transposing pre-pass)
        return launch_bp(p->fine, vol, vs, proj, ps, batch, ep, stream, wsp, ws_bytes, p->ss);
    }
    if (p->perm && !force_general()) {
        // the separable backprojector of the axis-swapped plan, then its
        // result moved into this plan's layout through the epilogue
        const int64_t ve = (int64_t)p->nx * p->ny * p->nz;
        const int64_t tmp_bytes = align_ws(ve * batch * 4), rest_bytes = workspace_bytes(p->perm, TSX_OP_BP, batch);
        Scratch owned;
        void *wptr = nullptr;
        int rc = get_workspace(p, tmp_bytes + rest_bytes, wsp, ws_bytes, stream, owned, &wptr);
        if (rc) return rc;
        float *tmp = (float *)wptr;
        rc = launch_bp(p->perm, tmp, ve, proj, ps, batch, plain(0), stream, (char *)wptr + tmp_bytes, rest_bytes,
                       up);
        if (rc) return rc;
        return launch_unswap(tmp, ve, vol, vs, p->nx, p->ny, p->nz, p->swap, batch, ep, stream);
    }
    const int64_t sino = (int64_t)p->rows * p->n_angles * p->cols;
    if (!p->bp_separable || force_general()) {
        if (up != 1) return fail(TSX_ERR_UNSUPPORTED, "refined-detector plan without the separable backprojector");
        if (!force_general() && bp_general_prescaled(p, batch)) {
            Scratch owned;
            void *wptr = nullptr;
            int rc = get_workspace(p, sino * batch * 4, wsp, ws_bytes, stream, owned, &wptr);
            if (rc) return rc;
            float *ly = (float *)wptr;
            launch_scale_by_length(p, proj, ps, ly, sino, batch, stream);
            return launch_bp_general<float, true>(p, vol, vs, ly, sino, batch, ep, stream);
        }
        return launch_bp_general<float, false>(p, vol, vs, proj, ps, batch, ep, stream);
    }
    // transposed, padded projections of the angles of one launch (all of them
    // unless they reach chunk_bytes: then ranges of sorted angles, each range's
    // sums added to the volume of the previous ones)
    const int kc = bp_chunk_of(p);
    const int64_t sino_t = bp_workspace_elems(p);
    const int parts = bp_parts_of(p);  // (> 1 only where kc covers every angle)
    const int64_t ve = (int64_t)p->nx * p->ny * p->nz;
    Scratch owned;
    void *wptr = nullptr;
    int rc = get_workspace(p, workspace_bytes(p, TSX_OP_BP, batch), wsp, ws_bytes, stream, owned, &wptr);
    if (rc) return rc;
    float *ly = (float *)wptr;
    float *partial = parts > 1 ? (float *)((char *)wptr + align_ws(sino_t * batch * 4)) : nullptr;
    const int BZ = bp_brick_z_of(p);
    const int64_t bricks = (int64_t)((p->nx + BR - 1) / BR) * ((p->ny + BR - 1) / BR) * ((p->nz + BZ - 1) / BZ);
    if (bricks * parts > max_blocks(BR * BR)) return fail(TSX_ERR_UNSUPPORTED, "volume too large for one launch");
    const int64_t bc = std::min<int64_t>(max_blocks(BR * BR) / (bricks * parts), 65535);  // entries per launch
    for (int k0 = 0; k0 < p->n_angles; k0 += kc) {
        const int kn = std::min(kc, p->n_angles - k0);
        const Epilogue e = k0 == 0 ? ep : later_chunk(ep);
        launch_scale_by_length_t(p, k0, kn, proj, ps, ly, sino_t, batch, stream, up);
        for (int64_t b0 = 0; b0 < batch; b0 += bc) {
            const int64_t nbat = std::min(bc, batch - b0);
            BpView v = bp_view(p, BZ, vol + b0 * vs, vs, ly + b0 * sino_t, (uint32_t)(sino_t * 4), (int)nbat,
                               ep_from(e, b0), k0, kn);
            if (parts > 1) {
                v.parts = parts;
                v.part_len = bp_part_len_of(p);
                v.partial = partial + b0 * parts * ve;
            }
            const int n_blocks = (int)(bricks * parts * nbat);
            int shift;
            const unsigned grid = xcd_grid(n_blocks, &shift);
            if (BZ == 32)
                hipLaunchKernelGGL(bp_wave<32>, dim3(grid), dim3(BR, BR), 0, stream, v, n_blocks, shift);
            else if (BZ == 24)
                hipLaunchKernelGGL(bp_wave<24>, dim3(grid), dim3(BR, BR), 0, stream, v, n_blocks, shift);
            else
                hipLaunchKernelGGL(bp_wave<16>, dim3(grid), dim3(BR, BR), 0, stream, v, n_blocks, shift);
            TSX_HIP_CHECK(hipGetLastError());
            if (parts > 1) {
                hipLaunchKernelGGL(bp_parts_reduce, dim3((unsigned)((ve + 255) / 256), (unsigned)nbat), dim3(256), 0,
                                   stream, v.partial, parts, ve, vol + b0 * vs, vs, ep_from(e, b0));
                TSX_HIP_CHECK(hipGetLastError());
            }
        }
    }
    return TSX_OK;
}


template <typename T>
int launch_general(const tsx_plan *p, bool forward, T *vol, int64_t vs, T *proj, int64_t ps,
                   int64_t batch, Epilogue ep, hipStream_t stream)
{
    if (batch <= 0) return TSX_OK;
    if (forward) return launch_fp_general<T>(p, (const T *)vol, vs, proj, ps, batch, ep, stream);
    return launch_bp_general<T, false>(p, vol, vs, (const T *)proj, ps, batch, ep, stream);
}

// ---------------------------------------------------------------------------
// fused launches over many instances
// ---------------------------------------------------------------------------
constexpr int64_t WS_ALIGN = 256;

int64_t align_up(int64_t v) { return (v + WS_ALIGN - 1) / WS_ALIGN * WS_ALIGN; }

int check_multi(const tsx_plan *const *plans, int n)
{
    if (!plans || n < 0) return fail(TSX_ERR_INVALID_ARGUMENT, "plans must not be NULL");
    for (int k = 0; k < n; ++k) {
        int rc = check_plan(plans[k]);
        if (rc) return rc;
        if (plans[k]->device != plans[0]->device)
            return fail(TSX_ERR_INVALID_ARGUMENT, "all plans of a fused launch must be on one device");
    }
    return TSX_OK;
}

// Scratch of a fused launch: each instance's workspace.
int64_t multi_workspace_bytes(const tsx_plan *const *plans, int n, int op)
{
    int64_t total = 0;
    for (int k = 0; k < n; ++k) total += align_up(workspace_bytes(plans[k], op, 1));
    return total;
}

// Fused launches: instances on the fast kernels with the same kernel choice
// share one launch (up to MAX_FUSED instances, as many blocks
// as the dispatch limit takes), their padded layouts (forward: one launch of
// transposes, one of copies) or transposed projections (backward: one
// launch) built before it; an instance on the general kernels, with
// detector supersampling, in slabs or angle ranges, or with more blocks than
// one launch takes, gets its own launch(es). eps[k] is the epilogue of
// instance k.
int multi_fp(const tsx_plan *const *plans, const float *const *vols, float *const *projs,
             const Epilogue *eps, int n, void *wsp, int64_t ws_bytes, hipStream_t stream)
{
    const int64_t need = multi_workspace_bytes(plans, n, TSX_OP_FP);
    Scratch owned;
    void *wptr = nullptr;
    if (need > 0) {
        int rc = get_workspace(plans[0], need, wsp, ws_bytes, stream, owned, &wptr);
        if (rc) return rc;
    }
    const bool fast = !force_general();
    FpMulti m;
    m.n = 0;
    m.begin[0] = 0;
    LayoutBatch layouts(stream);
    int variant = -1, by = 8;
    auto flush = [&]() -> int {
        if (m.n == 0) return TSX_OK;
        layouts.flush();
        const unsigned grid = xcd_grid(m.begin[m.n], &m.shift);
        pad_begin(m.begin, m.n);
        hipLaunchKernelGGL(fp_multi_kernel(variant, by), dim3(grid), dim3(FP_BX, by), 0, stream, m);
        TSX_HIP_CHECK(hipGetLastError());
        m.n = 0;
        return TSX_OK;
    };
    // each instance's workspace slice; the fused instances are taken by kernel
    // choice (march variant, tile rows), so that instances of one choice share
    // launches however they are ordered (every instance writes only its own
    // projections, so the order of the launches does not change any result)
    std::vector<int64_t> off(n + 1, 0);
    std::vector<std::pair<int, int>> keys;
    for (int k = 0; k < n; ++k) {
        if (!vols[k] || !projs[k]) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
        off[k + 1] = off[k] + align_up(workspace_bytes(plans[k], TSX_OP_FP, 1));
        const std::pair<int, int> key(fp_variant_of(plans[k]), fp_tile_rows_of(plans[k]));
        if (std::find(keys.begin(), keys.end(), key) == keys.end()) keys.push_back(key);
    }
    for (size_t pass = 0; pass < keys.size(); ++pass)
    for (int k = 0; k < n; ++k) {
        const tsx_plan *p = plans[k];
        const int64_t wb = off[k + 1] - off[k];
        const int64_t woff = off[k];
        const int pby = fp_tile_rows_of(p);
        const int along = p->fp_column_waves ? p->rows : p->cols, across = p->fp_column_waves ? p->cols : p->rows;
        const int tiles_u = (along + FP_BX - 1) / FP_BX;
        const int tiles_v = (across + pby - 1) / pby;
        const int64_t blocks = (int64_t)tiles_u * tiles_v * p->n_angles;
        const int64_t cap = max_blocks(FP_BX * pby);
        if (!(fast && p->fp_padded && p->ss == 1 && p->slabs.empty() && !fp_swapped(p) && !p->fp_wide) ||
            blocks > cap) {
            // this instance takes its own launch(es), in the first pass
            if (pass > 0) continue;
            int rc = launch_fp(p, vols[k], 0, projs[k], 0, 1, eps[k], stream, wptr ? (char *)wptr + woff : nullptr, wb);
            if (rc) return rc;
            continue;
        }
        if (fp_variant_of(p) != keys[pass].first || pby != keys[pass].second) continue;
        if (m.n > 0 && (fp_variant_of(p) != variant || pby != by || (int64_t)m.begin[m.n] + blocks > cap)) {
            int rc = flush();
            if (rc) return rc;
        }
        variant = fp_variant_of(p);
        by = pby;
        const int es = esize_of(variant);
        int64_t offs[3], per_entry;
        layout_offsets(p, es, offs, &per_entry);
        float *base = (float *)((char *)wptr + woff);
        layouts.add(p, es, vols[k], base, offs);
        FpView &v = m.v[m.n];
        v.angles = p->dev;
        v.n_angles = p->n_angles;
        v.a0 = 0;
        v.na = p->n_angles;
        v.rows = p->rows;
        v.cols = p->cols;
        v.tiles_u = tiles_u;
        v.tiles_v = tiles_v;
        v.shear_rows = fp_shear_rows_of(p);
        v.column_waves = p->fp_column_waves ? 1 : 0;
        v.batch = 1;
        for (int c = 0; c < 3; ++c) {
            v.L.ptr[c] = p->axis_count[c] > 0 ? base + offs[c] : nullptr;
            v.L.bytes[c] = (uint32_t)(p->layout_elems[c] * 4 * es);
        }
        v.L.stride = 0;
        v.proj = projs[k];
        v.proj_stride = 0;
        v.ep = eps[k];
        m.begin[m.n + 1] = m.begin[m.n] + (int)blocks;
        m.n += 1;
        if (m.n == MAX_FUSED) {
            int rc = flush();
            if (rc) return rc;
        }
    }
    TSX_HIP_CHECK(hipGetLastError());
    return flush();
}

int multi_bp(const tsx_plan *const *plans, float *const *vols, const float *const *projs,
             const Epilogue *eps, int n, void *wsp, int64_t ws_bytes, hipStream_t stream)
{
    const int64_t need = multi_workspace_bytes(plans, n, TSX_OP_BP);
    Scratch owned;
    void *wptr = nullptr;
    if (need > 0) {
        int rc = get_workspace(plans[0], need, wsp, ws_bytes, stream, owned, &wptr);
        if (rc) return rc;
    }
    const bool fast = !force_general();
    int BZ = 32;
    bool PM = false;  // the group's instances sum in angle parts
    BpMulti m;
    m.n = 0;
    m.begin[0] = 0;
    ScaleJobs sj;
    int sj_max = 0;  // blocks of the largest job in sj
    auto flush = [&]() -> int {
        if (m.n == 0) return TSX_OK;
        sj.n = m.n;
        pad_begin(m.begin, m.n);
        hipLaunchKernelGGL(scale_by_length_t_multi, dim3((unsigned)sj_max, (unsigned)sj.n), dim3(32, 8), 0, stream,
                           sj);
        sj_max = 0;
        const unsigned grid = xcd_grid(m.begin[m.n], &m.shift);
        auto kern = BZ == 32 ? (PM ? bp_wave_multi<32, true> : bp_wave_multi<32, false>)
                  : BZ == 24 ? (PM ? bp_wave_multi<24, true> : bp_wave_multi<24, false>)
                             : (PM ? bp_wave_multi<16, true> : bp_wave_multi<16, false>);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BR, BR), 0, stream, m);
        TSX_HIP_CHECK(hipGetLastError());
        m.n = 0;
        return TSX_OK;
    };
    // each instance's workspace slice; the fused instances are taken by kernel
    // choice (brick depth, angle parts), so that instances of one choice share
    // launches however they are ordered (every instance writes only its own
    // volume, so the order of the launches does not change any result)
    std::vector<int64_t> off(n + 1, 0);
    std::vector<std::pair<int, bool>> keys;
    for (int k = 0; k < n; ++k) {
        if (!vols[k] || !projs[k]) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
        off[k + 1] = off[k] + align_up(workspace_bytes(plans[k], TSX_OP_BP, 1));
        const std::pair<int, bool> key(bp_brick_z_of(plans[k]), bp_parts_of(plans[k]) > 1);
        if (std::find(keys.begin(), keys.end(), key) == keys.end()) keys.push_back(key);
    }
    for (size_t pass = 0; pass < keys.size(); ++pass)
    for (int k = 0; k < n; ++k) {
        const tsx_plan *p = plans[k];
        const int64_t wb = off[k + 1] - off[k];
        const int64_t woff = off[k];
        const int bz = bp_brick_z_of(p);
        const int64_t blocks = (int64_t)((p->nx + BR - 1) / BR) * ((p->ny + BR - 1) / BR) * ((p->nz + bz - 1) / bz);
        ScaleJob job;
        job.sorted = p->dev + p->n_angles;
        job.kn = p->n_angles;
        job.n_angles = p->n_angles;
        job.rows = p->rows;
        job.cols = p->cols;
        job.proj = projs[k];
        const int64_t tblocks = scale_job_blocks(job);
        const int64_t cap = max_blocks(BR * BR);
        if (!(fast && p->bp_separable && bp_chunk_of(p) == p->n_angles) || blocks > cap ||
            tblocks > max_blocks(256)) {
            if (pass > 0) continue;  // (its own launch, in the first pass)
            int rc = launch_bp(p, vols[k], 0, projs[k], 0, 1, eps[k], stream, wptr ? (char *)wptr + woff : nullptr, wb);
            if (rc) return rc;
            continue;
        }
        const int parts = bp_parts_of(p);
        if (bz != keys[pass].first || (parts > 1) != keys[pass].second) continue;
        if (m.n > 0 && (bz != BZ || (parts > 1) != PM || (int64_t)m.begin[m.n] + blocks > cap ||
                        std::max<int64_t>(sj_max, tblocks) * (m.n + 1) > max_blocks(256))) {
            int rc = flush();
            if (rc) return rc;
        }
        BZ = bz;
        PM = parts > 1;
        float *ly = (float *)((char *)wptr + woff);
        const int64_t sino_t = bp_workspace_elems(p);
        job.out = ly;
        job.blocks = (int)tblocks;
        sj.job[m.n] = job;
        sj_max = std::max(sj_max, (int)tblocks);
        m.v[m.n] = bp_view(p, BZ, vols[k], 0, ly, (uint32_t)(sino_t * 4), 1, eps[k], 0, p->n_angles);
        if (parts > 1) {
            m.v[m.n].parts = parts;
            m.v[m.n].part_len = bp_part_len_of(p);
        }
        m.begin[m.n + 1] = m.begin[m.n] + (int)blocks;
        m.n += 1;
        if (m.n == MAX_FUSED) {
            int rc = flush();
            if (rc) return rc;
        }
    }
    TSX_HIP_CHECK(hipGetLastError());
    return flush();
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
struct TuneChoice {
    int fp_variant, bp_bz, fp_by, fp_shear;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune_cache;

// Angle counts are keyed by the next power of two: how well the forward
// kernel's resident blocks share a tile's slab of the volume changes with the
// angle count only gradually, and a server with many angle counts then tunes
// a few shapes, not one per request.
int angle_bucket(int n)
{
    int b = 1;
    while (b < n && b < (1 << 30)) b <<= 1;
    return b;
}

std::string tune_key(const tsx_plan *p)
{
    return std::to_string(p->device) + ":" + std::to_string(p->host.empty() ? 0 : p->host[0].cone) +
           ":" + std::to_string(p->rows) + "x" + std::to_string(p->cols) + ":" + std::to_string(p->nx) +
           "x" + std::to_string(p->ny) + "x" + std::to_string(p->nz) + ":" + std::to_string(p->ss) + ":" +
           std::to_string((int)p->fp_padded) + std::to_string((int)p->fp_pair) +
           std::to_string((int)p->bp_separable) + ":" +
           std::to_string(angle_bucket(p->n_angles)) + ":" + std::to_string(p->fp_shear_rows) +
           (p->fp_column_waves ? "c" : "r");
}

// Mean time of `reps` launches after one warm-up launch, in milliseconds.
template <typename F>
int time_launches(F &&launch, hipStream_t stream, hipEvent_t e0, hipEvent_t e1, int reps, float *ms)
{
    int rc = launch();
    if (rc) return rc;
    TSX_HIP_CHECK(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; ++r) {
        rc = launch();
        if (rc) return rc;
    }
    TSX_HIP_CHECK(hipEventRecord(e1, stream));
    TSX_HIP_CHECK(hipEventSynchronize(e1));
    float t = 0.0f;
    TSX_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return TSX_OK;
}

// Owns the device allocations of one autotuning run.
struct TuneScratch {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void *vol = nullptr, *proj = nullptr, *band_proj = nullptr;
    AngleParams *angles = nullptr, *band_angles = nullptr;
    uint32_t *rt = nullptr;
    ~TuneScratch()
    {
        if (rt) (void)hipFree(rt);
        if (band_proj) (void)hipFree(band_proj);
        if (band_angles) (void)hipFree(band_angles);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (vol) (void)hipFree(vol);
        if (proj) (void)hipFree(proj);
        if (angles) (void)hipFree(angles);
    }
};

// Detector rows of the autotuner's tile-row trials: all of them, or as many
// as 2 GiB of projections hold.
int tile_band_rows(const tsx_plan *p)
{
    const int64_t per_row = (int64_t)p->n_angles * p->cols * 4;
    return (int)std::max<int64_t>(1, std::min<int64_t>(p->rows, ((int64_t)1 << 31) / per_row));
}

int autotune(tsx_plan *p, hipStream_t stream)
{
    const std::string key = tune_key(p);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            p->fp_variant = it->second.fp_variant;
            p->bp_bz = it->second.bp_bz;
            p->fp_by = it->second.fp_by;
            p->fp_shear = it->second.fp_shear;
            return TSX_OK;
        }
    }
    // Backprojector trials: a sub-plan of at most 64 of the plan's angles, in
    // four runs of 16 consecutive angles (the plan's own angular spacing, so that
    // the cache reuse between neighbouring angles is as in the full launch)
    // starting in the four quadrants of the scan.
    tsx_plan sub;
    sub.device = p->device;
    sub.kind = p->kind;
    sub.rows = p->rows;
    sub.cols = p->cols;
    sub.nx = p->nx;
    sub.ny = p->ny;
    sub.nz = p->nz;
    sub.ss = p->ss;
    const int N = p->n_angles;
    std::vector<int> pick;
    if (N <= 64) {
        for (int a = 0; a < N; ++a) pick.push_back(a);
    } else {
        for (int k = 0; k < 4; ++k)
            for (int i = 0; i < 16; ++i) pick.push_back((k * (N / 4) + i) % N);
    }
    for (int a : pick) {
        sub.host.push_back(p->host[a]);
        sub.axis_count[p->host[a].m] += 1;
    }
    sub.n_angles = (int32_t)sub.host.size();
    for (int c = 0; c < 3; ++c) sub.layout_elems[c] = p->layout_elems[c];
    sub.fp_padded = p->fp_padded;
    sub.fp_pair = p->fp_pair;
    sub.bp_separable = p->bp_separable;
    sub.bp_chunk = p->bp_chunk;
    TuneScratch t;
    const size_t vol_bytes = (size_t)p->nx * p->ny * p->nz * 4;
    const size_t proj_bytes = (size_t)p->rows * sub.n_angles * p->cols * 4;
    const size_t band_bytes = (size_t)std::max(tile_band_rows(p), std::min(p->rows, 64)) * p->n_angles * p->cols * 4;
    const std::vector<AngleParams> sub_table = device_angle_table(sub.host);
    if (hipMalloc(&t.angles, sizeof(AngleParams) * sub_table.size()) != hipSuccess ||
        hipMalloc(&t.band_angles, sizeof(AngleParams) * p->host.size()) != hipSuccess ||
        hipMalloc(&t.vol, vol_bytes) != hipSuccess || hipMalloc(&t.proj, proj_bytes) != hipSuccess ||
        hipMalloc(&t.band_proj, band_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(TSX_ERR_OUT_OF_MEMORY, "autotune: no device memory for the trial buffers");
    }
    sub.dev = t.angles;
    TSX_HIP_CHECK(hipMemcpyAsync(t.angles, sub_table.data(), sizeof(AngleParams) * sub_table.size(),
                                 hipMemcpyHostToDevice, stream));
    {
        int rc = build_range_table(&sub, stream, &t.rt);
        if (rc) return rc;
        sub.bp_rt = t.rt;
    }
    TSX_HIP_CHECK(hipMemsetAsync(t.vol, 0, vol_bytes, stream));
    TSX_HIP_CHECK(hipMemsetAsync(t.proj, 0, proj_bytes, stream));
    TSX_HIP_CHECK(hipEventCreate(&t.e0));
    TSX_HIP_CHECK(hipEventCreate(&t.e1));
    const Epilogue ep = plain(0);
    float *vol = (float *)t.vol, *proj = (float *)t.proj;
    TuneChoice best{-1, -1, -1, -1};
    // (a plan whose forward projection runs on its axis-swapped plan leaves the
    // choice to that plan's own tuning)
    const bool fp_choices =
        p->fp_padded && p->ss >= 1 && p->ss <= 4 && !force_general() && !fp_swapped(p) && !fp_refined(p);
    if (fp_choices) {
        // The forward kernel walks all angles of one detector tile before the
        // next tile, so its cache reuse depends on the angle count: the
        // candidates are timed on every angle of the plan, over bands of 64 and
        // 32 detector rows through the middle of the detector (the band's
        // pixel (0, 0) moved by r0 rows). A launch costs a fixed part (building
        // the padded layouts, larger for the row-paired ones) plus a part per
        // row; both are fitted from the two bands and the candidates compared
        // on the predicted time of the whole plan.
        // (a plan projected in z-slabs is timed on its first slab)
        const tsx_plan *fpp = p->slabs.empty() ? p : p->slabs[0];
        const int band_hi = std::min(p->rows, 64), band_lo = std::min(p->rows, 32);
        std::vector<AngleParams> shifted(fpp->host);
        tsx_plan band;
        band.device = p->device;
        band.kind = p->kind;
        band.cols = p->cols;
        band.nx = fpp->nx;
        band.ny = fpp->ny;
        band.nz = fpp->nz;
        band.ss = p->ss;
        band.n_angles = p->n_angles;
        for (int c = 0; c < 3; ++c) {
            band.axis_count[c] = fpp->axis_count[c];
            band.layout_elems[c] = fpp->layout_elems[c];
        }
        band.fp_padded = p->fp_padded;
        band.fp_pair = p->fp_pair;
        band.fp_shear_rows = fpp->fp_shear_rows;
        band.fp_column_waves = fpp->fp_column_waves;
        band.fp_wide = fpp->fp_wide;
        band.dev = t.band_angles;
        auto time_band = [&](int cand, int by, int band_rows, float *ms) -> int {
            const int r0 = (p->rows - band_rows) / 2;
            for (int a = 0; a < p->n_angles; ++a)
                for (int k = 0; k < 3; ++k) shifted[a].P0[k] = fpp->host[a].P0[k] + r0 * fpp->host[a].V[k];
            TSX_HIP_CHECK(hipMemcpyAsync(t.band_angles, shifted.data(), sizeof(AngleParams) * shifted.size(),
                                         hipMemcpyHostToDevice, stream));
            band.rows = band_rows;
            band.fp_variant = cand;
            band.fp_by = by;
            return time_launches([&] { return launch_fp(&band, vol, 0, (float *)t.band_proj, 0, 1, ep, stream); },
                                 stream, t.e0, t.e1, 2, ms);
        };
        auto predict = [&](int cand, int by, double *out) -> int {
            float ms_hi = 0.0f, ms_lo = 0.0f;
            int rc = time_band(cand, by, band_hi, &ms_hi);
            if (rc) return rc;
            *out = ms_hi;
            if (band_hi < p->rows) {
                rc = time_band(cand, by, band_lo, &ms_lo);
                if (rc) return rc;
                const double per_row = std::max(0.0, (double)(ms_hi - ms_lo) / (band_hi - band_lo));
                const double fixed = std::max(0.0, ms_hi - per_row * band_hi);
                *out = fixed + per_row * p->rows;
            }
            return TSX_OK;
        };
        // the march on tiles of 8 rows, then the tile rows of the chosen march
        double best_ms = 1e300;
        for (int cand : {4, 5, 3, 2}) {
            if (cand == 4 && !p->fp_pair) continue;
            if (p->fp_wide && (cand == 5 || cand == 2)) continue;
            double predicted = 0.0;
            int rc = predict(cand, 8, &predicted);
            if (rc) return rc;
            if (predicted < best_ms) {
                best_ms = predicted;
                best.fp_variant = cand;
                best.fp_by = 8;
            }
        }
        if (p->ss == 1 && !p->fp_wide && (best.fp_variant == 3 || best.fp_variant == 4)) {
            // Tile rows change how many blocks share a detector tile's volume
            // data and how the short rays near the detector's edges group;
            // they are timed on the whole plan (every angle and row, at most
            // 2 GiB of projections: a band through the middle otherwise).
            // Bands of 128 and 256 middle rows preferred 4-row tiles, 4 %
            // slower than 8-row ones on the whole 512^3 plan. 4-row tiles
            // replace the 8-row default only when more than 3 % faster.
            // (16-row tiles were a candidate: within 2 % of 8-row ones on one
            // 512^3 plan, ahead or behind from run to run, and 6 % slower in
            // the fused launch of 16 instances; their kernels stay, for
            // TSX_FP_TILE_ROWS.)
            const int band_tile = tile_band_rows(p);
            float tile_best = 1e30f;
            for (int by : {8, 4}) {
                float ms = 0.0f;
                int rc = time_band(best.fp_variant, by, band_tile, &ms);
                if (rc) return rc;
                if (by != 8) ms *= 1.03f;
                if (ms < tile_best) {
                    tile_best = ms;
                    best.fp_by = by;
                }
            }
        }
        if (fpp->fp_shear_rows > 0) {
            // Sheared tiles (a detector whose rows cross the slices' rows)
            // against tiles along the detector rows, on the whole plan: at
            // 256^3 the shear gained 27-56 % for detectors rolled by 0.5 rad
            // and between -4 and +10 % for 0.2 rad. Kept unless the plain
            // tiles are more than 2 % faster.
            const int by = best.fp_by > 0 ? best.fp_by : 8;
            float on_ms = 0.0f, off_ms = 0.0f;
            band.fp_shear = 1;
            int rc = time_band(best.fp_variant, by, tile_band_rows(p), &on_ms);
            band.fp_shear = 0;
            if (rc == TSX_OK) rc = time_band(best.fp_variant, by, tile_band_rows(p), &off_ms);
            band.fp_shear = -1;
            if (rc) return rc;
            best.fp_shear = off_ms * 1.02f < on_ms ? 0 : 1;
        }
    }
    if (p->bp_separable && !force_general()) {
        float best_ms = 1e30f;
        for (int bz : {32, 24, 16}) {
            sub.bp_bz = bz;
            float ms = 0.0f;
            int rc = time_launches([&] { return launch_bp(&sub, vol, 0, proj, 0, 1, ep, stream); },
                                   stream, t.e0, t.e1, 2, &ms);
            if (rc) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best.bp_bz = bz;
            }
        }
    }
    TSX_HIP_CHECK(hipStreamSynchronize(stream));
    p->fp_variant = best.fp_variant;
    p->bp_bz = best.bp_bz;
    p->fp_by = best.fp_by;
    p->fp_shear = best.fp_shear;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    return TSX_OK;
}

// A plan and the plans that run its kernels (axis-swapped, refined-detector).
int autotune_tree(tsx_plan *p, hipStream_t stream)
{
    int rc = autotune(p, stream);
    if (rc == TSX_OK && p->perm) rc = autotune_tree(p->perm, stream);
    if (rc == TSX_OK && p->fine) rc = autotune_tree(p->fine, stream);
    return rc;
}

bool tuned_tree(const tsx_plan *p)
{
    return p->fp_variant >= 0 || p->bp_bz > 0 || (p->perm && tuned_tree(p->perm)) || (p->fine && tuned_tree(p->fine));
}

// The z-slabs of a plan whose forward layouts reach chunk_bytes: plans of the
// same rays over slices [z0, z0 + depth) of the volume (the window's z range
// cut at the slab's faces), each with its angle table on `device` (none for
// a host-only plan). The parent's choice of row pairing holds for the slabs.
int build_slabs(const tsx_geometry *g, tsx_plan *p, int device)
{
    if (!p->fp_padded || p->fp_slab_depth >= p->nz) return TSX_OK;
    const double hz = (g->window_max[2] - g->window_min[2]) / (double)p->nz;
    for (int z0 = 0; z0 < p->nz; z0 += p->fp_slab_depth) {
        const int d = std::min(p->fp_slab_depth, p->nz - z0);
        tsx_geometry sg = *g;
        sg.grid_slice_count = d;
        sg.window_min[2] = g->window_min[2] + z0 * hz;
        sg.window_max[2] = z0 + d == p->nz ? g->window_max[2] : g->window_min[2] + (z0 + d) * hz;
        auto *s = new tsx_plan();
        p->slabs.push_back(s);
        p->slab_z0.push_back(z0);
        s->kind = p->kind;
        s->n_angles = p->n_angles;
        s->rows = p->rows;
        s->cols = p->cols;
        s->nx = p->nx;
        s->ny = p->ny;
        s->nz = d;
        s->ss = p->ss;
        s->chunk_bytes = p->chunk_bytes;
        s->is_slab = true;
        int rc = build_angles(&sg, s);
        if (rc) return rc;
        s->fp_padded = p->fp_padded;
        s->fp_pair = p->fp_pair;
        s->fdk_host.clear();
        if (device >= 0) {
            const std::vector<AngleParams> table = device_angle_table(s->host);
            const size_t bytes = sizeof(AngleParams) * table.size();
            hipError_t e = hipMalloc((void **)&s->dev, bytes);
            if (e != hipSuccess) {
                s->dev = nullptr;
                return fail(TSX_ERR_OUT_OF_MEMORY, std::string("slab angle table: ") + hipGetErrorString(e));
            }
            s->device = device;
            e = hipMemcpy(s->dev, table.data(), bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(TSX_ERR_HIP, std::string("slab angle table: ") + hipGetErrorString(e));
        }
    }
    p->slab_z0.push_back(p->nz);
    return TSX_OK;
}

// 1 if the detector's v axis is parallel to x at every angle, 2 if to y,
// else 0.
int v_parallel_swap(const tsx_geometry *g)
{
    for (int axis = 0; axis < 2; ++axis) {
        bool all = true;
        for (int a = 0; a < g->n_angles && all; ++a) {
            const double *v = g->vectors + 12 * (int64_t)a + 9;
            const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            for (int c = 0; c < 3; ++c)
                if (c != axis && !(std::fabs(v[c]) <= 1e-12 * vn)) all = false;
        }
        if (all) return axis + 1;
    }
    return 0;
}

// The geometry with volume axis swap - 1 (x or y) and z exchanged: grid
// counts, window and every vector's components; `vec` holds the vectors.
tsx_geometry swapped_geometry(const tsx_geometry *g, int swap, std::vector<double> &vec)
{
    const int a = swap - 1;
    tsx_geometry sg = *g;
    vec.assign(g->vectors, g->vectors + 12 * (int64_t)g->n_angles);
    for (int64_t k = 0; k < 4 * (int64_t)g->n_angles; ++k) std::swap(vec[3 * k + a], vec[3 * k + 2]);
    sg.vectors = vec.data();
    int32_t n[3] = {g->grid_col_count, g->grid_row_count, g->grid_slice_count};
    std::swap(n[a], n[2]);
    sg.grid_col_count = n[0];
    sg.grid_row_count = n[1];
    sg.grid_slice_count = n[2];
    std::swap(sg.window_min[a], sg.window_min[2]);
    std::swap(sg.window_max[a], sg.window_max[2]);
    return sg;
}

// The geometry on the detector refined ss times: rows and columns times ss,
// pixel vectors u / ss and v / ss about the same detector centre, no
// supersampling. Its pixel (i ss + p, j ss + q) is the sub-ray (p, q) of
// pixel (i, j): centre + (j + (q + 1/2) / ss - cols / 2) u + (same for rows).
tsx_geometry refined_geometry(const tsx_geometry *g, std::vector<double> &vec)
{
    const int ss = g->det_supersampling;
    tsx_geometry fg = *g;
    vec.assign(g->vectors, g->vectors + 12 * (int64_t)g->n_angles);
    for (int64_t a = 0; a < g->n_angles; ++a)
        for (int c = 6; c < 12; ++c) vec[12 * a + c] /= ss;
    fg.vectors = vec.data();
    fg.det_row_count = g->det_row_count * ss;
    fg.det_col_count = g->det_col_count * ss;
    fg.det_supersampling = 1;
    return fg;
}

// The upright FDK backprojector reads row-paired filtered projections (8 bytes
// per pixel, addressed with 32-bit offsets); the general one plain ones.
bool fdk_paired(const tsx_plan *p)
{
    return p->fdk_upright && !force_general() && p->cols >= 2 &&
           (int64_t)p->rows * p->n_angles * p->cols * 8 < ((int64_t)1 << 32);
}

}  // namespace

extern "C" {

const char *tsx_last_error(void) { return g_last_error.c_str(); }

const char *tsx_version(void) { return "tsx 0.1.0 (gfx950)"; }

int tsx_plan_create(const tsx_geometry *g, int device, tsx_plan **out)
{
    g_last_error.clear();
    if (!g || !out) return fail(TSX_ERR_INVALID_ARGUMENT, "geometry and out must not be NULL");
    *out = nullptr;
    if (g->type != TSX_PARALLEL3D_VEC && g->type != TSX_CONE_VEC)
        return fail(TSX_ERR_UNSUPPORTED, "projection type must be parallel3d_vec (0) or cone_vec (1)");
    if (g->n_angles <= 0 || !g->vectors)
        return fail(TSX_ERR_INVALID_ARGUMENT, "geometry needs at least one angle and a vectors array");
    if (g->det_row_count <= 0 || g->det_col_count <= 0)
        return fail(TSX_ERR_INVALID_ARGUMENT, "detector shape must be positive");
    if (g->grid_col_count <= 0 || g->grid_row_count <= 0 || g->grid_slice_count <= 0)
        return fail(TSX_ERR_INVALID_ARGUMENT, "volume shape must be positive");
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(g->window_min[c]) || !std::isfinite(g->window_max[c]) ||
            !(g->window_max[c] > g->window_min[c]))
            return fail(TSX_ERR_INVALID_ARGUMENT, "volume window must be finite with max > min");
    if (g->det_supersampling < 1 || g->det_supersampling > 16)
        return fail(TSX_ERR_INVALID_ARGUMENT, "detector supersampling must be in [1, 16]");
    if (g->vox_supersampling != 1)
        return fail(TSX_ERR_UNSUPPORTED,
                    "voxel supersampling != 1 is not supported: the backprojection is the exact "
                    "transpose of the forward projection");
    auto *p = new tsx_plan();
    p->kind = g->type;
    p->n_angles = g->n_angles;
    p->rows = g->det_row_count;
    p->cols = g->det_col_count;
    p->nx = g->grid_col_count;
    p->ny = g->grid_row_count;
    p->nz = g->grid_slice_count;
    p->ss = g->det_supersampling;
    p->chunk_bytes = chunk_limit();
    int rc = build_angles(g, p);
    if (rc != TSX_OK) {
        delete p;
        return rc;
    }
    if (device >= 0) {
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || device >= count) {
            delete p;
            return fail(TSX_ERR_NO_DEVICE, "device " + std::to_string(device) + " is not available");
        }
        DeviceGuard guard(device);
        if (!guard.ok) {
            delete p;
            return fail(TSX_ERR_HIP, "hipSetDevice failed");
        }
        const std::vector<AngleParams> table = device_angle_table(p->host);
        size_t bytes = sizeof(AngleParams) * table.size();
        e = hipMalloc((void **)&p->dev, bytes);
        if (e != hipSuccess) {
            delete p;
            return fail(TSX_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        e = hipMemcpy(p->dev, table.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(p->dev);
            delete p;
            return fail(TSX_ERR_HIP, std::string("hipMemcpy failed: ") + hipGetErrorString(e));
        }
        if (!p->fdk_host.empty()) {
            const size_t fb = sizeof(FdkAngle) * p->fdk_host.size();
            e = hipMalloc((void **)&p->fdk_dev, fb);
            if (e == hipSuccess) e = hipMemcpy(p->fdk_dev, p->fdk_host.data(), fb, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                if (p->fdk_dev) (void)hipFree(p->fdk_dev);
                (void)hipFree(p->dev);
                delete p;
                return fail(TSX_ERR_OUT_OF_MEMORY, std::string("FDK tables: ") + hipGetErrorString(e));
            }
        }
        int rc = build_range_table(p, nullptr, &p->bp_rt);
        if (rc == TSX_OK && hipStreamSynchronize(nullptr) != hipSuccess)
            rc = fail(TSX_ERR_HIP, "slice-range table: synchronisation failed");
        if (rc) {
            if (p->bp_rt) (void)hipFree(p->bp_rt);
            if (p->fdk_dev) (void)hipFree(p->fdk_dev);
            (void)hipFree(p->dev);
            delete p;
            return rc;
        }
        p->device = device;
    }
    if (device >= 0) {
        DeviceGuard guard(device);
        rc = build_slabs(g, p, device);
    } else {
        rc = build_slabs(g, p, -1);
    }
    if (rc) {
        const std::string msg = g_last_error;
        tsx_plan_destroy(p);
        g_last_error = msg;
        return rc;
    }
    // a detector whose v axis is parallel to x or y: the separable
    // backprojector of the same scan with that axis and z swapped
    const int swap = p->bp_separable || p->ss != 1 ? 0 : v_parallel_swap(g);
    if (swap) {
        std::vector<double> vec;
        tsx_geometry sg = swapped_geometry(g, swap, vec);
        tsx_plan *q = nullptr;
        if (tsx_plan_create(&sg, device, &q) == TSX_OK && q->bp_separable) {
            p->perm = q;
            p->swap = swap;
        } else {
            tsx_plan_destroy(q);
        }
        g_last_error.clear();
    }
    // detector supersampling: the same scan on the refined detector, for its
    // fast kernels
    if (p->ss > 1 && (int64_t)g->det_col_count * p->ss <= BP_RT_MAX_COLS &&
        (int64_t)g->det_row_count * p->ss <= ((int64_t)1 << 24)) {
        std::vector<double> vec;
        tsx_geometry fg = refined_geometry(g, vec);
        tsx_plan *q = nullptr;
        if (tsx_plan_create(&fg, device, &q) == TSX_OK &&
            (q->bp_separable || q->perm || q->fp_padded || (q->perm && q->perm->fp_padded))) {
            p->fine = q;
        } else {
            tsx_plan_destroy(q);
        }
        g_last_error.clear();
    }
    *out = p;
    return TSX_OK;
}

int tsx_plan_destroy(tsx_plan *p)
{
    if (!p) return TSX_OK;
    for (tsx_plan *s : p->slabs) tsx_plan_destroy(s);
    tsx_plan_destroy(p->perm);
    tsx_plan_destroy(p->fine);
    if (p->dev) {
        DeviceGuard guard(p->device);
        (void)hipFree(p->dev);
        if (p->fdk_dev) (void)hipFree(p->fdk_dev);
        if (p->bp_rt) (void)hipFree(p->bp_rt);
    }
    delete p;
    return TSX_OK;
}

int tsx_plan_info_get(const tsx_plan *p, tsx_plan_info *info)
{
    if (!p || !info) return fail(TSX_ERR_INVALID_ARGUMENT, "plan and info must not be NULL");
    info->device = p->device;
    info->n_angles = p->n_angles;
    info->det_rows = p->rows;
    info->det_cols = p->cols;
    info->vol_shape[0] = p->nz;
    info->vol_shape[1] = p->ny;
    info->vol_shape[2] = p->nx;
    tsx_plan_info sub;
    int64_t sub_bytes = p->perm && tsx_plan_info_get(p->perm, &sub) == TSX_OK ? sub.plan_bytes : 0;
    if (p->fine && tsx_plan_info_get(p->fine, &sub) == TSX_OK) sub_bytes += sub.plan_bytes;
    info->plan_bytes = sub_bytes + (p->dev ? (int64_t)(sizeof(AngleParams) * 2 * p->host.size()) *
                                        (1 + (int64_t)p->slabs.size()) +
                                    (p->bp_rt ? p->bp_rt_stride * p->n_angles * 4 : 0)
                              : 0);
    for (int c = 0; c < 3; ++c) info->axis_count[c] = p->axis_count[c];
    // the plan whose forward kernels run
    const tsx_plan *f0 = fp_refined(p) ? p->fine : p;
    const tsx_plan *f = fp_swapped(f0) ? f0->perm : f0;
    info->fp_kernel = f->fp_padded && f->ss <= 4 && !force_general() ? 1 : 0;
    const tsx_plan *b = bp_refined(p) ? p->fine : p;  // the plan whose backprojector runs,
    const tsx_plan *bs = b->perm ? b->perm : b;                      // on its axis-swapped plan if any
    info->bp_kernel = (b->bp_separable || b->perm) && !force_general() ? 1 : 0;
    info->fp_variant = fp_variant_of(f);
    info->bp_brick_z = bp_brick_z_of(bs);
    info->tuned = tuned_tree(p);
    info->fp_slabs = info->fp_kernel ? std::max<int32_t>(1, (int32_t)f->slabs.size()) : 0;
    info->fp_tile_rows = info->fp_kernel ? fp_tile_rows_of(f) : 0;
    info->fp_axis_swap = fp_swapped(f0) ? f0->swap : 0;
    info->fp_refined_detector = f0 != p ? p->ss : 0;
    info->fp_shear_rows = info->fp_kernel ? fp_shear_rows_of(f) : 0;
    info->fp_column_waves = info->fp_kernel && f->fp_column_waves ? 1 : 0;
    info->bp_launch_angles = info->bp_kernel ? bp_chunk_of(bs) : 0;
    info->bp_axis_swap = info->bp_kernel ? b->swap : 0;
    info->bp_refined_detector = b != p ? p->ss : 0;
    return TSX_OK;
}

int tsx_fdk_workspace_size(const tsx_plan *p, int64_t *bytes)
{
    if (!p || !bytes) return fail(TSX_ERR_INVALID_ARGUMENT, "plan and bytes must not be NULL");
    *bytes = (int64_t)p->rows * p->n_angles * p->cols * (fdk_paired(p) ? 8 : 4);
    return TSX_OK;
}

int tsx_fdk(const tsx_plan *p, float *vol, const float *proj, int additive, void *workspace,
            int64_t workspace_bytes, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    if (p->kind != TSX_CONE_VEC || !p->fdk_dev)
        return fail(TSX_ERR_UNSUPPORTED, "FDK needs a cone-beam (cone_vec) geometry");
    if (p->ss != 1)
        return fail(TSX_ERR_UNSUPPORTED, "FDK with detector supersampling is not supported");
    if (p->cols > FDK_MAX_COLS)
        return fail(TSX_ERR_UNSUPPORTED, "FDK supports at most " + std::to_string(FDK_MAX_COLS) +
                                             " detector columns");
    if (p->rows > 65535 || p->n_angles > 65535 || p->nz > 65535)
        return fail(TSX_ERR_UNSUPPORTED, "FDK launch grid limit (65535 rows, angles or slices)");
    TSX_ENTRY("tsx_fdk");
    DeviceGuard dg(p->device);
    if (!dg.ok) return fail(TSX_ERR_HIP, "cannot select the plan's device");
    EntryTimer entry_timer_("tsx_fdk", stream);
    hipStream_t s = (hipStream_t)stream;
    const bool upright = fdk_paired(p);
    const int64_t need = (int64_t)p->rows * p->n_angles * p->cols * (upright ? 8 : 4);
    Scratch owned;
    void *wptr = nullptr;
    rc = get_workspace(p, need, workspace, workspace_bytes, s, owned, &wptr);
    if (rc) return rc;
    float *q = (float *)wptr;
    const size_t lds = (size_t)fdk_lds_floats(p->cols) * sizeof(float);
    double tail = 0.25;  // 0.25 - 2 * sum of the filter's lag coefficients
    for (int k = fdk_nodd_pad(p->cols) - 1; k >= 0; --k) {
        const double m = 2.0 * k + 1.0;
        tail -= 2.0 / (M_PI * M_PI * m * m);
    }
    hipLaunchKernelGGL(fdk_filter, dim3((unsigned)p->rows, (unsigned)p->n_angles), dim3(FDK_FILTER_THREADS),
                       lds, s, p->fdk_dev, p->n_angles, p->rows, p->cols, proj, q, upright ? 1 : 0,
                       (float)tail);
    TSX_HIP_CHECK(hipGetLastError());
    const unsigned gx = (unsigned)((p->nx + FDK_BX - 1) / FDK_BX), gy = (unsigned)((p->ny + FDK_BY - 1) / FDK_BY);
    if (upright)
        hipLaunchKernelGGL(fdk_bp_upright, dim3(gx, gy, (unsigned)((p->nz + FDK_ZC - 1) / FDK_ZC)),
                           dim3(FDK_BX, FDK_BY), 0, s, p->fdk_dev, p->n_angles, p->rows, p->cols, q, vol,
                           p->nx, p->ny, p->nz, p->vlo[0], p->vlo[1], p->vlo[2], p->vh[0], p->vh[1],
                           p->vh[2], additive ? 1 : 0);
    else
        hipLaunchKernelGGL(fdk_bp, dim3(gx, gy, (unsigned)p->nz), dim3(FDK_BX, FDK_BY), 0, s, p->fdk_dev,
                           p->n_angles, p->rows, p->cols, q, vol, p->nx, p->ny, p->nz, p->vlo[0],
                           p->vlo[1], p->vlo[2], p->vh[0], p->vh[1], p->vh[2], additive ? 1 : 0);
    TSX_HIP_CHECK(hipGetLastError());
    return TSX_OK;
}

int tsx_plan_autotune(tsx_plan *p, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    TraceRange trace("tsx_plan_autotune");
    DeviceGuard dg(p->device);
    if (!dg.ok) return fail(TSX_ERR_HIP, "cannot select the plan's device");
    rc = autotune_tree(p, (hipStream_t)stream);
    // the trials' stream-ordered scratch (freed, the stream synchronised) goes
    // back to the device instead of staying in the default pool, which the
    // library keeps from shrinking, for the rest of the process
    hipMemPool_t pool;
    if (hipStreamSynchronize((hipStream_t)stream) == hipSuccess &&
        hipDeviceGetDefaultMemPool(&pool, p->device) == hipSuccess)
        (void)hipMemPoolTrimTo(pool, 0);
    return rc;
}

int tsx_plan_axes(const tsx_plan *p, int32_t *axes)
{
    if (!p || !axes) return fail(TSX_ERR_INVALID_ARGUMENT, "plan and axes must not be NULL");
    for (int a = 0; a < p->n_angles; ++a) axes[a] = p->host[a].m;
    return TSX_OK;
}

int tsx_workspace_size(const tsx_plan *p, int op, int64_t batch, int64_t *bytes)
{
    if (!p || !bytes) return fail(TSX_ERR_INVALID_ARGUMENT, "plan and bytes must not be NULL");
    if (op != TSX_OP_FP && op != TSX_OP_BP)
        return fail(TSX_ERR_INVALID_ARGUMENT, "op must be TSX_OP_FP (0) or TSX_OP_BP (1)");
    *bytes = workspace_bytes(p, op, batch < 1 ? 1 : batch);
    return TSX_OK;
}

int tsx_forward_batched(const tsx_plan *p, const float *vol, int64_t vs, float *proj, int64_t ps,
                        int64_t batch, int additive, void *workspace, int64_t workspace_bytes,
                        void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    TSX_ENTRY("tsx_forward_batched");
    DeviceGuard guard(p->device);
    EntryTimer entry_timer_("tsx_forward_batched", stream);
    return launch_fp(p, vol, vs, proj, ps, batch, plain(additive), (hipStream_t)stream, workspace,
                     workspace_bytes);
}

int tsx_backward_batched(const tsx_plan *p, float *vol, int64_t vs, const float *proj, int64_t ps,
                         int64_t batch, int additive, void *workspace, int64_t workspace_bytes,
                         void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    TSX_ENTRY("tsx_backward_batched");
    DeviceGuard guard(p->device);
    EntryTimer entry_timer_("tsx_backward_batched", stream);
    return launch_bp(p, vol, vs, proj, ps, batch, plain(additive), (hipStream_t)stream, workspace,
                     workspace_bytes);
}

int tsx_forward_batched_f64(const tsx_plan *p, const double *vol, int64_t vs, double *proj,
                            int64_t ps, int64_t batch, int additive, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    TSX_ENTRY("tsx_forward_batched_f64");
    DeviceGuard guard(p->device);
    EntryTimer entry_timer_("tsx_forward_batched_f64", stream);
    return launch_general<double>(p, true, (double *)vol, vs, proj, ps, batch, plain(additive),
                                  (hipStream_t)stream);
}

int tsx_backward_batched_f64(const tsx_plan *p, double *vol, int64_t vs, const double *proj,
                             int64_t ps, int64_t batch, int additive, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    TSX_ENTRY("tsx_backward_batched_f64");
    DeviceGuard guard(p->device);
    EntryTimer entry_timer_("tsx_backward_batched_f64", stream);
    return launch_general<double>(p, false, vol, vs, (double *)proj, ps, batch, plain(additive),
                                  (hipStream_t)stream);
}

int tsx_forward_residual(const tsx_plan *p, const float *vol, int64_t vs, const float *y,
                         const float *w, float *proj, int64_t ps, int64_t batch, void *workspace,
                         int64_t workspace_bytes, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj || !y || !w)
        return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    Epilogue e;
    e.mode = EP_RESIDUAL;
    e.y = y;
    e.w = w;
    e.stride = ps;
    TSX_ENTRY("tsx_forward_residual");
    DeviceGuard guard(p->device);
    EntryTimer entry_timer_("tsx_forward_residual", stream);
    return launch_fp(p, vol, vs, proj, ps, batch, e, (hipStream_t)stream, workspace,
                     workspace_bytes);
}

int tsx_backward_scaled_add(const tsx_plan *p, float *vol, int64_t vs, const float *proj,
                            int64_t ps, const float *w, int64_t batch, void *workspace,
                            int64_t workspace_bytes, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj || !w) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    Epilogue e;
    e.mode = EP_SCALED_ADD;
    e.y = nullptr;
    e.w = w;
    e.stride = vs;
    TSX_ENTRY("tsx_backward_scaled_add");
    DeviceGuard guard(p->device);
    EntryTimer entry_timer_("tsx_backward_scaled_add", stream);
    return launch_bp(p, vol, vs, proj, ps, batch, e, (hipStream_t)stream, workspace,
                     workspace_bytes);
}

int tsx_multi_workspace_size(const tsx_plan *const *plans, int n, int op, int64_t *bytes)
{
    if (!plans || !bytes || n < 0) return fail(TSX_ERR_INVALID_ARGUMENT, "plans and bytes must not be NULL");
    if (op != TSX_OP_FP && op != TSX_OP_BP)
        return fail(TSX_ERR_INVALID_ARGUMENT, "op must be TSX_OP_FP (0) or TSX_OP_BP (1)");
    for (int k = 0; k < n; ++k)
        if (!plans[k]) return fail(TSX_ERR_INVALID_ARGUMENT, "plan is NULL");
    *bytes = multi_workspace_bytes(plans, n, op);
    return TSX_OK;
}

}  // extern "C"

namespace {

// Per-instance epilogues of a fused call: SET / ADD for all when `eps` is
// NULL, else as given (checked).
int multi_epilogues(const tsx_epilogue *eps, int n, int additive, int op, std::vector<Epilogue> &out)
{
    out.resize(n);
    for (int k = 0; k < n; ++k) {
        Epilogue &e = out[k];
        e.y = nullptr;
        e.w = nullptr;
        e.stride = 0;
        if (!eps) {
            e.mode = additive ? EP_ADD : EP_SET;
            continue;
        }
        const int mode = eps[k].mode;
        if (mode == TSX_EP_SET || mode == TSX_EP_ADD) {
            e.mode = mode == TSX_EP_SET ? EP_SET : EP_ADD;
        } else if (mode == TSX_EP_RESIDUAL && op == TSX_OP_FP) {
            if (!eps[k].y || !eps[k].w)
                return fail(TSX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(k) +
                                                          ": the residual epilogue needs y and w");
            e.mode = EP_RESIDUAL;
            e.y = eps[k].y;
            e.w = eps[k].w;
        } else if (mode == TSX_EP_SCALED_ADD) {
            if (!eps[k].w)
                return fail(TSX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(k) +
                                                          ": the scaled-add epilogue needs w");
            e.mode = EP_SCALED_ADD;
            e.w = eps[k].w;
        } else {
            return fail(TSX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(k) + ": epilogue mode " +
                                                      std::to_string(mode) + " is not valid for this call");
        }
    }
    return TSX_OK;
}

int multi_entry(const tsx_plan *const *plans, const void *const *vols, const void *const *projs,
                const tsx_epilogue *eps, int n, int additive, int op, void *workspace,
                int64_t workspace_bytes, void *stream)
{
    int rc = check_multi(plans, n);
    if (rc) return rc;
    if (n == 0) return TSX_OK;
    if (!vols || !projs) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointer arrays must not be NULL");
    std::vector<Epilogue> ep;
    rc = multi_epilogues(eps, n, additive, op, ep);
    if (rc) return rc;
    TSX_ENTRY(op == TSX_OP_FP ? "tsx_multi_forward" : "tsx_multi_backward");
    DeviceGuard guard(plans[0]->device);
    EntryTimer entry_timer_(op == TSX_OP_FP ? "tsx_multi_forward" : "tsx_multi_backward", stream);
    if (op == TSX_OP_FP)
        return multi_fp(plans, (const float *const *)vols, (float *const *)projs, ep.data(), n, workspace,
                        workspace_bytes, (hipStream_t)stream);
    return multi_bp(plans, (float *const *)vols, (const float *const *)projs, ep.data(), n, workspace,
                    workspace_bytes, (hipStream_t)stream);
}

}  // namespace

extern "C" {

int tsx_multi_forward(const tsx_plan *const *plans, const float *const *vols, float *const *projs,
                      int n, int additive, void *workspace, int64_t workspace_bytes, void *stream)
{
    return multi_entry(plans, (const void *const *)vols, (const void *const *)projs, nullptr, n, additive,
                       TSX_OP_FP, workspace, workspace_bytes, stream);
}

int tsx_multi_backward(const tsx_plan *const *plans, float *const *vols, const float *const *projs,
                       int n, int additive, void *workspace, int64_t workspace_bytes, void *stream)
{
    return multi_entry(plans, (const void *const *)vols, (const void *const *)projs, nullptr, n, additive,
                       TSX_OP_BP, workspace, workspace_bytes, stream);
}

int tsx_multi_forward_ep(const tsx_plan *const *plans, const float *const *vols, float *const *projs,
                         const tsx_epilogue *eps, int n, void *workspace, int64_t workspace_bytes,
                         void *stream)
{
    if (!eps && n > 0) return fail(TSX_ERR_INVALID_ARGUMENT, "epilogues must not be NULL");
    return multi_entry(plans, (const void *const *)vols, (const void *const *)projs, eps, n, 0, TSX_OP_FP,
                       workspace, workspace_bytes, stream);
}

int tsx_multi_backward_ep(const tsx_plan *const *plans, float *const *vols, const float *const *projs,
                          const tsx_epilogue *eps, int n, void *workspace, int64_t workspace_bytes,
                          void *stream)
{
    if (!eps && n > 0) return fail(TSX_ERR_INVALID_ARGUMENT, "epilogues must not be NULL");
    return multi_entry(plans, (const void *const *)vols, (const void *const *)projs, eps, n, 0, TSX_OP_BP,
                       workspace, workspace_bytes, stream);
}

int tsx_forward(const tsx_plan *p, const float *vol, float *proj, int additive, void *stream)
{
    return tsx_forward_batched(p, vol, 0, proj, 0, 1, additive, nullptr, 0, stream);
}

int tsx_backward(const tsx_plan *p, float *vol, const float *proj, int additive, void *stream)
{
    return tsx_backward_batched(p, vol, 0, proj, 0, 1, additive, nullptr, 0, stream);
}

int tsx_direct_fpbp3d(const tsx_plan *p, float *vol, float *proj, int mode, int op, void *stream)
{
    if (mode != TSX_MODE_SET && mode != TSX_MODE_ADD)
        return fail(TSX_ERR_INVALID_ARGUMENT, "mode must be TSX_MODE_SET (1) or TSX_MODE_ADD (0)");
    const int additive = mode == TSX_MODE_ADD;
    if (op == TSX_OP_FP) return tsx_forward(p, vol, proj, additive, stream);
    if (op == TSX_OP_BP) return tsx_backward(p, vol, proj, additive, stream);
    return fail(TSX_ERR_INVALID_ARGUMENT, "op must be TSX_OP_FP (0) or TSX_OP_BP (1)");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// No-allocation variant of tsx_direct_fpbp3d. The kernels' workspace (padded
// volume layouts for the forward projector, length-weighted projections for
// the backprojector) lives in `scratch`, a static device array that every
// GPU running this library holds until the process ends. Its host-side name
// is only a stand-in: host code uses the device address that
// hipGetSymbolAddress returns, once per device with that device current.
// Every byte of the workspace that a kernel reads is written earlier in the
// same call (the layout builders write the zero borders too, the weighting
// pass every projection element), so nothing is cleared between calls.
// Calls that overlap on one device share the array and race.
static __device__ __align__(256) unsigned char scratch[TSX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

// Device address of `scratch` on `device` (which must be current).
int scratch_address(int device, void **out)
{
    static std::mutex mu;
    static void *addr[MAX_DEVICES] = {nullptr};
    if (device < 0 || device >= MAX_DEVICES)
        return fail(TSX_ERR_UNSUPPORTED, "device index " + std::to_string(device) + " is beyond " +
                                             std::to_string(MAX_DEVICES - 1));
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(TSX_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        addr[device] = ptr;
    }
    *out = addr[device];
    return TSX_OK;
}

}  // namespace

extern "C" {

int tsx_direct_fpbp3d_noalloc(const tsx_plan *p, float *vol, float *proj, int mode, int op, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (mode != TSX_MODE_SET && mode != TSX_MODE_ADD)
        return fail(TSX_ERR_INVALID_ARGUMENT, "mode must be TSX_MODE_SET (1) or TSX_MODE_ADD (0)");
    if (op != TSX_OP_FP && op != TSX_OP_BP)
        return fail(TSX_ERR_INVALID_ARGUMENT, "op must be TSX_OP_FP (0) or TSX_OP_BP (1)");
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    const int64_t need = workspace_bytes(p, op, 1);
    if (need > TSX_NOALLOC_SCRATCH_BYTES)
        return fail(TSX_ERR_UNSUPPORTED, "tsx_direct_fpbp3d_noalloc: the plan needs " + std::to_string(need) +
                                             " bytes of workspace, more than the " +
                                             std::to_string((int64_t)TSX_NOALLOC_SCRATCH_BYTES) +
                                             " bytes of the static scratch buffer");
    TSX_ENTRY("tsx_direct_fpbp3d_noalloc");
    DeviceGuard guard(p->device);
    if (!guard.ok) return fail(TSX_ERR_HIP, "cannot select the plan's device");
    EntryTimer entry_timer_("tsx_direct_fpbp3d_noalloc", stream);
    void *ws = nullptr;
    rc = scratch_address(p->device, &ws);
    if (rc) return rc;
    const Epilogue ep = plain(mode == TSX_MODE_ADD);
    if (op == TSX_OP_FP)
        return launch_fp(p, vol, 0, proj, 0, 1, ep, (hipStream_t)stream, ws, TSX_NOALLOC_SCRATCH_BYTES);
    return launch_bp(p, vol, 0, proj, 0, 1, ep, (hipStream_t)stream, ws, TSX_NOALLOC_SCRATCH_BYTES);
}

}  // extern "C"
