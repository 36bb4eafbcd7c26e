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

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tsx.h"
#include "tsx_plan.h"

using tsx::AngleParams;

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

struct tsx_plan {
    int device = -1;
    int32_t kind = 0;
    int32_t n_angles = 0, rows = 0, cols = 0;
    int32_t nx = 0, ny = 0, nz = 0;
    int32_t ss = 1;  // detector supersampling
    std::vector<AngleParams> host;
    AngleParams *dev = nullptr;
    int32_t axis_count[3] = {0, 0, 0};
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

// Dominant axis: argmax |e_c|, ties resolved in the order x, y, z.
int dominant_axis(V3 e)
{
    double ax = std::fabs(e[0]), ay = std::fabs(e[1]), az = std::fabs(e[2]);
    if (ax >= ay && ax >= az) return 0;
    if (ay >= az) return 1;
    return 2;
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
        A.kappa = cone ? dot(P0 - S, N) : 0.0;
        A.EN = cone ? 0.0 : dot(S, N);
        A.m = m;
        A.c1 = c1;
        A.c2 = c2;
        A.cone = cone ? 1 : 0;
        p->axis_count[m] += 1;
    }
    return TSX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace {

constexpr int FP_BX = 64, FP_BY = 4;
constexpr int BP_BX = 64, BP_BY = 4;

__device__ __forceinline__ int sel3(int axis, int x, int y, int z)
{
    return axis == 0 ? x : (axis == 1 ? y : z);
}

// One (sub-)ray of the forward projector. (fj, fi) are detector coordinates in
// pixel units with pixel centres on integers.
__device__ float fp_ray(const AngleParams &A, const float *__restrict__ vol, double fj, double fi)
{
    const double Pm = A.P0[0] + fj * A.U[0] + fi * A.V[0];
    const double P1 = A.P0[1] + fj * A.U[1] + fi * A.V[1];
    const double P2 = A.P0[2] + fj * A.U[2] + fi * A.V[2];
    const double Dm = A.cone ? Pm - A.S[0] : A.S[0];
    const double D1 = A.cone ? P1 - A.S[1] : A.S[1];
    const double D2 = A.cone ? P2 - A.S[2] : A.S[2];
    if (Dm == 0.0) return 0.0f;
    const double inv = 1.0 / Dm;
    const double a1 = D1 * inv, a2 = D2 * inv;
    const double len = sqrt(A.h[0] * A.h[0] + A.h[1] * A.h[1] * a1 * a1 + A.h[2] * A.h[2] * a2 * a2);
    const int nm = A.n[0], n1 = A.n[1], n2 = A.n[2];
    const int nref = nm / 2;
    const double R1 = P1 + ((double)nref - Pm) * a1;
    const double R2 = P2 + ((double)nref - Pm) * a2;
    // slices where (Q1, Q2) can touch the volume: Q_c in (-1, n_c)
    double lo = -1e30, hi = 1e30;
    if (a1 != 0.0) {
        double t0 = (-1.0 - R1) / a1, t1 = ((double)n1 - R1) / a1;
        lo = fmax(lo, fmin(t0, t1));
        hi = fmin(hi, fmax(t0, t1));
    } else if (!(R1 > -1.0 && R1 < (double)n1)) {
        return 0.0f;
    }
    if (a2 != 0.0) {
        double t0 = (-1.0 - R2) / a2, t1 = ((double)n2 - R2) / a2;
        lo = fmax(lo, fmin(t0, t1));
        hi = fmin(hi, fmax(t0, t1));
    } else if (!(R2 > -1.0 && R2 < (double)n2)) {
        return 0.0f;
    }
    lo = fmax(lo + nref, -1.0);
    hi = fmin(hi + nref, (double)nm);
    if (!(lo <= hi)) return 0.0f;
    const int nlo = max(0, (int)floor(lo));
    const int nhi = min(nm - 1, (int)ceil(hi));
    const float fa1 = (float)a1, fa2 = (float)a2, fR1 = (float)R1, fR2 = (float)R2;
    const int64_t sm = A.st[0], s1 = A.st[1], s2 = A.st[2];
    float total = 0.0f;
    float part = 0.0f;
    int cnt = 0;
    for (int n = nlo; n <= nhi; ++n) {
        const float dn = (float)(n - nref);
        const float q1 = fmaf(dn, fa1, fR1);
        const float q2 = fmaf(dn, fa2, fR2);
        const float f1 = floorf(q1), f2 = floorf(q2);
        const int k1 = (int)f1, k2 = (int)f2;
        const float w1 = q1 - f1, w2 = q2 - f2;
        const bool a0 = (unsigned)k1 < (unsigned)n1, a1ok = (unsigned)(k1 + 1) < (unsigned)n1;
        const bool b0 = (unsigned)k2 < (unsigned)n2, b1ok = (unsigned)(k2 + 1) < (unsigned)n2;
        const int64_t base = (int64_t)n * sm + (int64_t)k1 * s1 + (int64_t)k2 * s2;
        const float v00 = (a0 && b0) ? vol[base] : 0.0f;
        const float v10 = (a1ok && b0) ? vol[base + s1] : 0.0f;
        const float v01 = (a0 && b1ok) ? vol[base + s2] : 0.0f;
        const float v11 = (a1ok && b1ok) ? vol[base + s1 + s2] : 0.0f;
        const float lo2 = fmaf(w1, v10 - v00, v00);
        const float hi2 = fmaf(w1, v11 - v01, v01);
        part += fmaf(w2, hi2 - lo2, lo2);
        if (++cnt == 32) {
            total += part;
            part = 0.0f;
            cnt = 0;
        }
    }
    total += part;
    return (float)len * total;
}

__global__ void __launch_bounds__(FP_BX *FP_BY)
fp_kernel(const AngleParams *__restrict__ angles, int n_angles, int rows, int cols, int ss,
          int tiles_u, const float *__restrict__ vol, int64_t vol_stride, float *__restrict__ proj,
          int64_t proj_stride, int additive)
{
    const int tile = blockIdx.x;
    const int j = (tile % tiles_u) * FP_BX + threadIdx.x;
    const int i = (tile / tiles_u) * FP_BY + threadIdx.y;
    const int a = blockIdx.y;
    const int b = blockIdx.z;
    if (j >= cols || i >= rows) return;
    const AngleParams &A = angles[a];
    const float *v = vol + (int64_t)b * vol_stride;
    float acc;
    if (ss == 1) {
        acc = fp_ray(A, v, (double)j, (double)i);
    } else {
        acc = 0.0f;
        const double inv = 1.0 / ss;
        for (int p = 0; p < ss; ++p)
            for (int q = 0; q < ss; ++q)
                acc += fp_ray(A, v, j + (q + 0.5) * inv - 0.5, i + (p + 0.5) * inv - 0.5);
        acc *= (float)(inv * inv);
    }
    float *out = proj + (int64_t)b * proj_stride + ((int64_t)i * n_angles + a) * cols + j;
    *out = additive ? *out + acc : acc;
}

// Exact transpose of fp_kernel, voxel-driven. For every angle the voxel's
// slice n = k_m is intersected by the rays whose samples in that slice lie
// within the open square |Q - k| < 1. The rays through the voxel centre meet
// the detector at (cj, ci); around it the in-slice offset of a ray at
// detector offset (dj, di) is the projective map
//   dq_c = (dj*A_cj + di*A_ci) / (1 + dj*Bj + di*Bi)   (Bj = Bi = 0 for parallel),
// whose inverse images of the square corners bound the pixels to visit.
__global__ void __launch_bounds__(BP_BX *BP_BY)
bp_kernel(const AngleParams *__restrict__ angles, int n_angles, int rows, int cols, int ss,
          int nx, int ny, int nz, float *__restrict__ vol, int64_t vol_stride,
          const float *__restrict__ proj, int64_t proj_stride, int additive)
{
    const int x = blockIdx.x * BP_BX + threadIdx.x;
    const int y = blockIdx.y * BP_BY + threadIdx.y;
    const int z = blockIdx.z % nz;
    const int b = blockIdx.z / nz;
    if (x >= nx || y >= ny) return;
    const float *pr = proj + (int64_t)b * proj_stride;
    const float inv_ss = 1.0f / (float)ss;
    const float sub_margin = 0.5f - 0.5f * inv_ss;
    float acc = 0.0f;
    for (int a = 0; a < n_angles; ++a) {
        const AngleParams &A = angles[a];
        const double km = sel3(A.m, x, y, z), k1 = sel3(A.c1, x, y, z), k2 = sel3(A.c2, x, y, z);
        double cj, ci, r1, r2, T, Bjc = 0.0, Bic = 0.0, g = 0.0;
        if (A.cone) {
            const double W0 = km - A.S[0], W1 = k1 - A.S[1], W2 = k2 - A.S[2];
            const double dn = W0 * A.N[0] + W1 * A.N[1] + W2 * A.N[2];
            if (dn == 0.0 || W0 == 0.0) continue;
            const double lam = A.kappa / dn;
            const double R0 = A.S[0] + lam * W0 - A.P0[0];
            const double R1 = A.S[1] + lam * W1 - A.P0[1];
            const double R2 = A.S[2] + lam * W2 - A.P0[2];
            cj = R0 * A.Bj[0] + R1 * A.Bj[1] + R2 * A.Bj[2];
            ci = R0 * A.Bi[0] + R1 * A.Bi[1] + R2 * A.Bi[2];
            r1 = W1 / W0;
            r2 = W2 / W0;
            T = 1.0 / lam;
            Bjc = A.U[0] / (lam * W0);
            Bic = A.V[0] / (lam * W0);
            g = 1.0 / W0;
        } else {
            const double t0 = ((A.P0[0] - km) * A.N[0] + (A.P0[1] - k1) * A.N[1] +
                               (A.P0[2] - k2) * A.N[2]) / A.EN;
            const double R0 = km + t0 * A.S[0] - A.P0[0];
            const double R1 = k1 + t0 * A.S[1] - A.P0[1];
            const double R2 = k2 + t0 * A.S[2] - A.P0[2];
            cj = R0 * A.Bj[0] + R1 * A.Bj[1] + R2 * A.Bj[2];
            ci = R0 * A.Bi[0] + R1 * A.Bi[1] + R2 * A.Bi[2];
            r1 = A.S[1] / A.S[0];
            r2 = A.S[2] / A.S[0];
            T = 1.0;
        }
        const float A1j = (float)(T * (A.U[1] - r1 * A.U[0]));
        const float A1i = (float)(T * (A.V[1] - r1 * A.V[0]));
        const float A2j = (float)(T * (A.U[2] - r2 * A.U[0]));
        const float A2i = (float)(T * (A.V[2] - r2 * A.V[0]));
        const float fBj = (float)Bjc, fBi = (float)Bic;
        // inverse map of the square corners (dq1, dq2) = (+-1, +-1)
        const float det = A1j * A2i - A1i * A2j;
        if (det == 0.0f) continue;
        const float idet = 1.0f / det;
        const float M00 = A2i * idet, M01 = -A1i * idet, M10 = -A2j * idet, M11 = A1j * idet;
        const float cv0 = fBj * M00 + fBi * M10, cv1 = fBj * M01 + fBi * M11;
        float djmin = 1e30f, djmax = -1e30f, dimin = 1e30f, dimax = -1e30f;
        bool unbounded = false;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float s1 = (s & 1) ? 1.0f : -1.0f, s2 = (s & 2) ? 1.0f : -1.0f;
            const float den = 1.0f - (cv0 * s1 + cv1 * s2);
            if (!(den > 1e-6f)) unbounded = true;
            const float iden = 1.0f / den;
            const float dj = (M00 * s1 + M01 * s2) * iden, di = (M10 * s1 + M11 * s2) * iden;
            djmin = fminf(djmin, dj);
            djmax = fmaxf(djmax, dj);
            dimin = fminf(dimin, di);
            dimax = fmaxf(dimax, di);
        }
        int j0, j1, i0, i1;
        if (unbounded) {
            j0 = 0; j1 = cols - 1; i0 = 0; i1 = rows - 1;
        } else {
            const double mj = sub_margin + 1e-3, mi = sub_margin + 1e-3;
            j0 = (int)fmax(0.0, ceil(cj + djmin - mj));
            j1 = (int)fmin((double)(cols - 1), floor(cj + djmax + mj));
            i0 = (int)fmax(0.0, ceil(ci + dimin - mi));
            i1 = (int)fmin((double)(rows - 1), floor(ci + dimax + mi));
        }
        if (j0 > j1 || i0 > i1) continue;
        const float fr1 = (float)r1, fr2 = (float)r2, fg = (float)g;
        const float h0 = (float)(A.h[0] * A.h[0]), h1 = (float)(A.h[1] * A.h[1]),
                    h2 = (float)(A.h[2] * A.h[2]);
        const float len_par = sqrtf(h0 + h1 * fr1 * fr1 + h2 * fr2 * fr2);
        const float dj0 = (float)((double)j0 - cj), di0 = (float)((double)i0 - ci);
        float acc_a = 0.0f;
        for (int i = i0; i <= i1; ++i) {
            const float* row = pr + ((int64_t)i * n_angles + a) * cols;
            for (int j = j0; j <= j1; ++j) {
                const float yv = row[j];
                float wsum = 0.0f;
                for (int p = 0; p < ss; ++p) {
                    const float di = di0 + (float)(i - i0) + ((p + 0.5f) * inv_ss - 0.5f);
                    for (int q = 0; q < ss; ++q) {
                        const float dj = dj0 + (float)(j - j0) + ((q + 0.5f) * inv_ss - 0.5f);
                        const float den = fmaf(dj, fBj, fmaf(di, fBi, 1.0f));
                        const float iden = 1.0f / den;
                        const float dq1 = fmaf(dj, A1j, di * A1i) * iden;
                        const float dq2 = fmaf(dj, A2j, di * A2i) * iden;
                        const float w = fmaxf(0.0f, 1.0f - fabsf(dq1)) * fmaxf(0.0f, 1.0f - fabsf(dq2));
                        float len = len_par;
                        if (A.cone) {
                            const float s1 = fmaf(dq1, fg, fr1), s2 = fmaf(dq2, fg, fr2);
                            len = sqrtf(h0 + h1 * s1 * s1 + h2 * s2 * s2);
                        }
                        wsum = fmaf(w, len, wsum);
                    }
                }
                acc_a = fmaf(wsum, yv, acc_a);
            }
        }
        acc += acc_a * (inv_ss * inv_ss);
    }
    float *out = vol + (int64_t)b * vol_stride + ((int64_t)z * ny + y) * nx + x;
    *out = additive ? *out + acc : acc;
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

int launch_fp(const tsx_plan *p, const float *vol, int64_t vs, float *proj, int64_t ps,
              int64_t batch, int additive, hipStream_t stream)
{
    if (batch <= 0) return TSX_OK;
    if (p->n_angles > 65535 || batch > 65535)
        return fail(TSX_ERR_UNSUPPORTED, "more than 65535 angles or batch elements per launch");
    const int tiles_u = (p->cols + FP_BX - 1) / FP_BX;
    const int tiles_v = (p->rows + FP_BY - 1) / FP_BY;
    dim3 grid((unsigned)(tiles_u * tiles_v), (unsigned)p->n_angles, (unsigned)batch);
    dim3 block(FP_BX, FP_BY, 1);
    hipLaunchKernelGGL(fp_kernel, grid, block, 0, stream, p->dev, p->n_angles, p->rows, p->cols,
                       p->ss, tiles_u, vol, vs, proj, ps, additive);
    TSX_HIP_CHECK(hipGetLastError());
    return TSX_OK;
}

int launch_bp(const tsx_plan *p, float *vol, int64_t vs, const float *proj, int64_t ps,
              int64_t batch, int additive, hipStream_t stream)
{
    if (batch <= 0) return TSX_OK;
    if ((int64_t)p->nz * batch > 65535)
        return fail(TSX_ERR_UNSUPPORTED, "volume depth times batch exceeds 65535 per launch");
    dim3 grid((unsigned)((p->nx + BP_BX - 1) / BP_BX), (unsigned)((p->ny + BP_BY - 1) / BP_BY),
              (unsigned)(p->nz * batch));
    dim3 block(BP_BX, BP_BY, 1);
    hipLaunchKernelGGL(bp_kernel, grid, block, 0, stream, p->dev, p->n_angles, p->rows, p->cols,
                       p->ss, p->nx, p->ny, p->nz, vol, vs, proj, ps, additive);
    TSX_HIP_CHECK(hipGetLastError());
    return TSX_OK;
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
        size_t bytes = sizeof(AngleParams) * p->host.size();
        e = hipMalloc((void **)&p->dev, bytes);
        if (e != hipSuccess) {
            delete p;
            return fail(TSX_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        e = hipMemcpy(p->dev, p->host.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(p->dev);
            delete p;
            return fail(TSX_ERR_HIP, std::string("hipMemcpy failed: ") + hipGetErrorString(e));
        }
        p->device = device;
    }
    *out = p;
    return TSX_OK;
}

int tsx_plan_destroy(tsx_plan *p)
{
    if (!p) return TSX_OK;
    if (p->dev) {
        DeviceGuard guard(p->device);
        (void)hipFree(p->dev);
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
    info->plan_bytes = p->dev ? (int64_t)(sizeof(AngleParams) * p->host.size()) : 0;
    for (int c = 0; c < 3; ++c) info->axis_count[c] = p->axis_count[c];
    return TSX_OK;
}

int tsx_plan_axes(const tsx_plan *p, int32_t *axes)
{
    if (!p || !axes) return fail(TSX_ERR_INVALID_ARGUMENT, "plan and axes must not be NULL");
    for (int a = 0; a < p->n_angles; ++a) axes[a] = p->host[a].m;
    return TSX_OK;
}

int tsx_forward_batched(const tsx_plan *p, const float *vol, int64_t vs, float *proj, int64_t ps,
                        int64_t batch, int additive, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    DeviceGuard guard(p->device);
    return launch_fp(p, vol, vs, proj, ps, batch, additive, (hipStream_t)stream);
}

int tsx_backward_batched(const tsx_plan *p, float *vol, int64_t vs, const float *proj, int64_t ps,
                         int64_t batch, int additive, void *stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!vol || !proj) return fail(TSX_ERR_INVALID_ARGUMENT, "data pointers must not be NULL");
    DeviceGuard guard(p->device);
    return launch_bp(p, vol, vs, proj, ps, batch, additive, (hipStream_t)stream);
}

int tsx_forward(const tsx_plan *p, const float *vol, float *proj, int additive, void *stream)
{
    return tsx_forward_batched(p, vol, 0, proj, 0, 1, additive, stream);
}

int tsx_backward(const tsx_plan *p, float *vol, const float *proj, int additive, void *stream)
{
    return tsx_backward_batched(p, vol, 0, proj, 0, 1, additive, stream);
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
