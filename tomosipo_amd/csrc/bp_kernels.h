// Backprojection kernels (included by tsx.hip). Both compute the exact
// transpose of the Joseph forward projector:
//   x[k] = sum_{a, i, j} l(a,i,j) * L(Q_c1 - k_c1) * L(Q_c2 - k_c2) * y[i, a, j],
// where Q is the intersection of ray (a, i, j) with slice n = k_m,
// L(t) = max(0, 1 - |t|) and l is the world length per slice step.
//
// bp_general: one thread per voxel; for every angle it maps the voxel's
//   +-1 square in its slice to the detector (projective map around the
//   voxel-centre ray) and sums over the pixels in the bounding box. Any
//   geometry, any detector supersampling.
//
// bp_separable: for geometries whose detector v axis is parallel to the
//   in-slice axis c1 = z and whose stepping axis is x or y (every circular,
//   helical or jittered-source scan with an upright detector). Then Q_c2 of a
//   ray depends only on its detector column j, and Q_z is affine in the row i,
//   so per slice n
//     x[n, kz, kc2] = sum_j L(Q_c2(j; n) - kc2) * u[n, j, kz],
//     u[n, j, kz]   = sum_i L(A(j, n) + i B(j, n) - kz) * ly[i, j],
//   with ly = l * y precomputed. A block owns a brick of 16 x 16 columns by
//   BZ voxels along z; per angle it builds u for its 16 slices in LDS (pass
//   1: thread (slice, column) splats its detector column onto the z-grid,
//   rows in increasing z with the pending cells kept in registers, so that
//   every cell is written without being read back), stores the columns of a
//   slice in increasing order of Q_c2, and then every thread folds exactly
//   the columns with |Q_c2 - kc2| < 1 into the BZ voxels of its z-column
//   (pass 2; the first one found by binary search). The slice ranges of the
//   next angle are computed during the current one. No atomics: the result
//   is deterministic.
#pragma once

#include "tsx_plan.h"
#include "fp_kernels.h"

namespace tsx {

constexpr int BP_BX = 64, BP_BY = 4;

constexpr int BR = 16;  // brick edge (x, y) of bp_separable

__device__ __forceinline__ int sel3(int axis, int x, int y, int z)
{
    return axis == 0 ? x : (axis == 1 ? y : z);
}

// Arithmetic of bp_general by precision: the float kernel takes the FP64
// reciprocal above for its FP64 setup and the hardware FP32 reciprocal and
// square root (1 ulp) per pixel; the double kernel (gradient checks) keeps
// exact divisions and square roots.
template <typename T>
struct GeneralMath;
template <>
struct GeneralMath<float> {
    static __device__ __forceinline__ double recip64(double x) { return rcp64(x); }
    static __device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float root(float x) { return __builtin_amdgcn_sqrtf(x); }
};
template <>
struct GeneralMath<double> {
    static __device__ __forceinline__ double recip64(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double recip(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double root(double x) { return sqrt(x); }
};

// With PRESCALED (detector supersampling 1 only), proj holds the
// length-weighted projections l(a, i, j) * y[i, a, j] and the per-pixel ray
// length is not recomputed.
template <typename T, bool PRESCALED = false>
__global__ void __launch_bounds__(BP_BX *BP_BY)
bp_general(const AngleParams *__restrict__ angles, int n_angles, int rows, int cols, int ss,
           int nx, int ny, int nz, T *__restrict__ vol, int64_t vol_stride,
           const T *__restrict__ proj, int64_t proj_stride, Epilogue ep)
{
    const int x = blockIdx.x * BP_BX + threadIdx.x;
    const int y = blockIdx.y * BP_BY + threadIdx.y;
    const int z = blockIdx.z % nz;
    const int b = blockIdx.z / nz;
    if (x >= nx || y >= ny) return;
    const T *pr = proj + (int64_t)b * proj_stride;
    const double sub_margin = 0.5 - 0.5 / ss;
    T acc = T(0);
    for (int a = 0; a < n_angles; ++a) {
        const AngleParams &A = angles[a];
        const double km = sel3(A.m, x, y, z), k1 = sel3(A.c1, x, y, z), k2 = sel3(A.c2, x, y, z);
        // (cj, ci): the voxel centre's detector position, FP64; the ray
        // slopes r1, r2 and the homography terms are O(1) and use T
        using M = GeneralMath<T>;
        double cj, ci;
        T r1, r2, Tc, Bjc = T(0), Bic = T(0), g = T(0);
        if (A.cone) {
            const double W0 = km - A.S[0], W1 = k1 - A.S[1], W2 = k2 - A.S[2];
            const double dn = W0 * A.N[0] + W1 * A.N[1] + W2 * A.N[2];
            if (dn == 0.0 || W0 == 0.0) continue;
            const double lam = A.kappa * M::recip64(dn);
            const double R0 = A.S[0] + lam * W0 - A.P0[0];
            const double R1 = A.S[1] + lam * W1 - A.P0[1];
            const double R2 = A.S[2] + lam * W2 - A.P0[2];
            cj = R0 * A.Bj[0] + R1 * A.Bj[1] + R2 * A.Bj[2];
            ci = R0 * A.Bi[0] + R1 * A.Bi[1] + R2 * A.Bi[2];
            const T iW0 = M::recip((T)W0);
            r1 = (T)W1 * iW0;
            r2 = (T)W2 * iW0;
            Tc = (T)dn * (T)M::recip64(A.kappa);  // 1 / lam
            Bjc = (T)A.U[0] * Tc * iW0;
            Bic = (T)A.V[0] * Tc * iW0;
            g = iW0;
        } else {
            const double t0 = ((A.P0[0] - km) * A.N[0] + (A.P0[1] - k1) * A.N[1] +
                               (A.P0[2] - k2) * A.N[2]) * M::recip64(A.EN);
            const double R0 = km + t0 * A.S[0] - A.P0[0];
            const double R1 = k1 + t0 * A.S[1] - A.P0[1];
            const double R2 = k2 + t0 * A.S[2] - A.P0[2];
            cj = R0 * A.Bj[0] + R1 * A.Bj[1] + R2 * A.Bj[2];
            ci = R0 * A.Bi[0] + R1 * A.Bi[1] + R2 * A.Bi[2];
            const double iS0 = M::recip64(A.S[0]);
            r1 = (T)(A.S[1] * iS0);
            r2 = (T)(A.S[2] * iS0);
            Tc = T(1);
        }
        const T A1j = Tc * ((T)A.U[1] - r1 * (T)A.U[0]);
        const T A1i = Tc * ((T)A.V[1] - r1 * (T)A.V[0]);
        const T A2j = Tc * ((T)A.U[2] - r2 * (T)A.U[0]);
        const T A2i = Tc * ((T)A.V[2] - r2 * (T)A.V[0]);
        const T fBj = (T)Bjc, fBi = (T)Bic;
        // the local homography (offsets of O(1) pixels and voxels) in the
        // kernel's precision; only the voxel's own detector position (cj, ci)
        // needs FP64
        const T det = A1j * A2i - A1i * A2j;
        if (det == T(0)) continue;
        const T idet = GeneralMath<T>::recip(det);
        const T M00 = A2i * idet, M01 = -A1i * idet, M10 = -A2j * idet, M11 = A1j * idet;
        const T cv0 = fBj * M00 + fBi * M10, cv1 = fBj * M01 + fBi * M11;
        T djmin = T(1e30), djmax = T(-1e30), dimin = T(1e30), dimax = T(-1e30);
        bool unbounded = false;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const T s1 = (s & 1) ? T(1) : T(-1), s2 = (s & 2) ? T(1) : T(-1);
            const T den = T(1) - (cv0 * s1 + cv1 * s2);
            if (!(den > T(1e-6))) unbounded = true;
            const T iden = GeneralMath<T>::recip(den);
            const T dj = (M00 * s1 + M01 * s2) * iden, di = (M10 * s1 + M11 * s2) * iden;
            djmin = fmin(djmin, dj);
            djmax = fmax(djmax, dj);
            dimin = fmin(dimin, di);
            dimax = fmax(dimax, di);
        }
        int j0, j1, i0, i1;
        if (unbounded) {
            j0 = 0; j1 = cols - 1; i0 = 0; i1 = rows - 1;
        } else {
            const double mj = sub_margin + 1e-3, mi = sub_margin + 1e-3;
            j0 = (int)fmax(0.0, ceil(cj + (double)djmin - mj));
            j1 = (int)fmin((double)(cols - 1), floor(cj + (double)djmax + mj));
            i0 = (int)fmax(0.0, ceil(ci + (double)dimin - mi));
            i1 = (int)fmin((double)(rows - 1), floor(ci + (double)dimax + mi));
        }
        if (j0 > j1 || i0 > i1) continue;
        const T fr1 = r1, fr2 = r2, fg = g;
        const T h0 = (T)(A.h[0] * A.h[0]), h1 = (T)(A.h[1] * A.h[1]), h2 = (T)(A.h[2] * A.h[2]);
        const T len_par = GeneralMath<T>::root(h0 + h1 * fr1 * fr1 + h2 * fr2 * fr2);
        const T dj0 = (T)((double)j0 - cj), di0 = (T)((double)i0 - ci);
        const T tinv = (T)1 / (T)ss;
        T acc_a = T(0);
        for (int i = i0; i <= i1; ++i) {
            const T *row = pr + ((int64_t)i * n_angles + a) * cols;
            for (int j = j0; j <= j1; ++j) {
                const T yv = row[j];
                T wsum = T(0);
                for (int p = 0; p < ss; ++p) {
                    const T di = di0 + (T)(i - i0) + ((T)(p + 0.5) * tinv - (T)0.5);
                    for (int q = 0; q < ss; ++q) {
                        const T dj = dj0 + (T)(j - j0) + ((T)(q + 0.5) * tinv - (T)0.5);
                        const T den = dj * fBj + (di * fBi + T(1));
                        const T iden = GeneralMath<T>::recip(den);
                        const T dq1 = (dj * A1j + di * A1i) * iden;
                        const T dq2 = (dj * A2j + di * A2i) * iden;
                        const T w1 = T(1) - fabs(dq1), w2 = T(1) - fabs(dq2);
                        const T w = (w1 > T(0) ? w1 : T(0)) * (w2 > T(0) ? w2 : T(0));
                        if constexpr (PRESCALED) {
                            wsum += w;
                        } else {
                            T len = len_par;
                            if (A.cone) {
                                const T s1 = dq1 * fg + fr1, s2 = dq2 * fg + fr2;
                                len = GeneralMath<T>::root(h0 + h1 * s1 * s1 + h2 * s2 * s2);
                            }
                            wsum += w * len;
                        }
                    }
                }
                acc_a += wsum * yv;
            }
        }
        acc += acc_a * (tinv * tinv);
    }
    emit(vol + (int64_t)b * vol_stride, ((int64_t)z * ny + y) * nx + x, b, acc, ep);
}

// ly[i, a, j] = l(a, i, j) * y[i, a, j] (detector supersampling 1). Block
// (x, y) takes columns 256 x .. 256 x + 255 of the detector lines y,
// y + gridDim.y, ... of the `lines` = batch * rows * n_angles lines (line
// (b, i, a) is b * rows * n_angles + i * n_angles + a).
__global__ void scale_by_length(const AngleParams *__restrict__ angles, int n_angles, int rows,
                                int cols, const float *__restrict__ proj, int64_t proj_stride,
                                float *__restrict__ out, int64_t out_stride, int64_t lines)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const int64_t per_entry = (int64_t)rows * n_angles;
    for (int64_t line = blockIdx.y; line < lines; line += gridDim.y) {
        const int64_t b = line / per_entry;
        const int64_t ia = line - b * per_entry;
        const int i = (int)(ia / n_angles), a = (int)(ia - (int64_t)i * n_angles);
        const AngleParams &A = angles[a];
        float D[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float P = (float)(A.P0[c] + (double)j * A.U[c] + (double)i * A.V[c]);
            D[c] = A.cone ? P - (float)A.S[c] : (float)A.S[c];
        }
        const float inv = 1.0f / D[0];
        const float a1 = D[1] * inv, a2 = D[2] * inv;
        const float len = (D[0] == 0.0f) ? 0.0f
            : sqrtf((float)(A.h[0] * A.h[0]) + (float)(A.h[1] * A.h[1]) * a1 * a1 +
                    (float)(A.h[2] * A.h[2]) * a2 * a2);
        const int64_t o = ia * cols + j;
        out[b * out_stride + o] = len * proj[b * proj_stride + o];
    }
}

// Detector column where the ray through in-slice point (n, c) (axes m, c2)
// meets the detector. Independent of z for an upright detector.
template <typename AP>
__device__ __forceinline__ double column_of(const AP &A, double n, double c)
{
    if (A.cone) {
        const double W0 = n - A.S[0], W2 = c - A.S[2];
        const double lam = A.kappa * rcp64(W0 * A.N[0] + W2 * A.N[2]);
        return (A.S[0] + lam * W0 - A.P0[0]) * A.Bj[0] + (A.S[2] + lam * W2 - A.P0[2]) * A.Bj[2];
    }
    const double t0 = ((A.P0[0] - n) * A.N[0] + (A.P0[2] - c) * A.N[2]) * rcp64(A.EN);
    return (n + t0 * A.S[0] - A.P0[0]) * A.Bj[0] + (c + t0 * A.S[2] - A.P0[2]) * A.Bj[2];
}

// One instance of the separable backprojector as seen by a block.
struct BpView {
    const AngleParams *angles;
    int n_angles, rows, cols, nx, ny, nz, bricks_x, bricks_y, bricks_z, batch;
    float *vol;
    int64_t vol_stride;
    const float *ly;    // length-weighted projections, batch entries ly_bytes apart
    uint32_t ly_bytes;
    Epilogue ep;
};

// LDS layout of bp_separable: columns are UCOL = BZ + 2 floats apart and
// slices SLICE floats apart (SLICE = 36 mod 64); the cell kz of the column in
// slot pos of slice s is at s * SLICE + pos * UCOL + UOFF + kz. Only the cells
// kz = 0 .. BZ - 1 are read. The splat clamps the lower cell of a row to
// [-2, BZ], so the weights of rows outside the brick land in cells -2, -1 of
// their column or in cells BZ, BZ + 1; a column's cells BZ - 2, BZ - 1 share
// the next column's (unused) cells -4, -3, and its cells BZ, BZ + 1 the next
// column's cells -2, -1, so no real cell is written by another column. The
// strides were chosen with a simulation of the kernel's LDS addresses on a
// cone-beam geometry (64 dword banks): they spread the z-cells that the lanes
// of a pass-1 wave (16 columns of 4 slices) and of a pass-2 wave (16 slices)
// touch at the same time over the banks (7 % fewer bank cycles than columns
// BZ + 4 apart), and every column stays 8-byte aligned so that pass 2 reads
// its cells as float2.
constexpr int UOFF = 4;  // kz = 0 sits at offset 4 of a column

template <int BZ>
struct BpShared {
    static constexpr int UCOL = BZ + 2;
    static constexpr int SLICE = (BR * UCOL + 27) / 64 * 64 + 36;  // first stride >= BR * UCOL + 2, = 36 mod 64
    static_assert(UCOL % 2 == 0 && SLICE % 64 == 36 && SLICE >= BR * UCOL + 2, "LDS strides");
    float u[BR * SLICE];
    float q2[BR][20];     // Q_c2 - c0 per stored column, increasing; 1e30 if unused
    int jlo[2][BR];       // first detector column of each slice (double-buffered by angle)
    int cnt[2][BR];       // number of columns of each slice
    int inc[2][BR];       // 1 if Q_c2 increases with the detector column
    int chunks[2];        // max over the slices of ceil(cnt / BR)
};


// Slice range of slice t of a brick (first column, count, direction).
template <typename AP>
__device__ __forceinline__ void bp_slice_range(const AP &A, int bx, int by, int cols, int t,
                                                int &jlo, int &cnt, int &inc)
{
    const bool mx = A.m == 0;
    const int m0 = mx ? bx * BR : by * BR;
    const int c0 = mx ? by * BR : bx * BR;
    const double n = (double)(m0 + t);
    const double ja = column_of(A, n, (double)c0 - 1.0);
    const double jb = column_of(A, n, (double)(c0 + BR));
    const int lo = max(0, (int)ceil(fmin(ja, jb) - 1e-3));
    const int hi = min(cols - 1, (int)floor(fmax(ja, jb) + 1e-3));
    jlo = lo;
    cnt = max(0, hi - lo + 1);
    inc = ja <= jb;
}

// Slice ranges of one angle for the 16 slices of a brick, computed by lanes
// t = 0 .. BR - 1 of the first wave (all active together), which also publish
// the number of column chunks.
template <int BZ, typename AP>
__device__ __forceinline__ void bp_ranges(const AP &A, int bx, int by, int cols, int t,
                                          BpShared<BZ> &sh, int buf)
{
    int jlo, cnt, inc;
    bp_slice_range(A, bx, by, cols, t, jlo, cnt, inc);
    sh.jlo[buf][t] = jlo;
    sh.cnt[buf][t] = cnt;
    sh.inc[buf][t] = inc;
    int ch = (cnt + BR - 1) / BR;
#pragma unroll
    for (int o = BR / 2; o > 0; o >>= 1) ch = max(ch, __shfl_xor(ch, o, BR));
    if (t == 0) sh.chunks[buf] = ch;
}

template <int BZ>
__device__ __forceinline__ void bp_separable_work(const BpView &v, int logical, BpShared<BZ> &sh)
{
    using S = BpShared<BZ>;
    const int n_angles = v.n_angles, rows = v.rows, cols = v.cols;
    const int bx = logical % v.bricks_x;
    const int by = (logical / v.bricks_x) % v.bricks_y;
    const int bz = (logical / (v.bricks_x * v.bricks_y)) % v.bricks_z;
    const int b = logical / (v.bricks_x * v.bricks_y * v.bricks_z);
    if (b >= v.batch) return;
    // The angle parameters do not change during the launch. Read through a
    // constant-address-space pointer, the uniform ones become scalar loads
    // into SGPRs; through a generic pointer the barriers of the angle loop
    // (memory fences) make the compiler read them with vector loads into
    // VGPRs (7 % slower at 32-deep bricks, 22 % at 16-deep ones).
    typedef const __attribute__((address_space(4))) AngleParams ConstAngle;
    ConstAngle *__restrict__ angles = (ConstAngle *)v.angles;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = bx * BR + tx, y = by * BR + ty, z0 = bz * BZ;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)uniform_ptr(v.ly + (int64_t)b * ((int64_t)v.ly_bytes / 4)), (short)0,
        (int)__builtin_amdgcn_readfirstlane(v.ly_bytes), 0x00020000);
    const uint32_t row_bytes = (uint32_t)n_angles * (uint32_t)cols * 4u;
    float acc[BZ];
#pragma unroll
    for (int k = 0; k < BZ; ++k) acc[k] = 0.0f;
    // slice ranges (16 lanes of the first wave, one slice each) of angle 0
    const int t = ty * BR + tx;
    if (t < BR) bp_ranges(angles[0], bx, by, cols, t, sh, 0);
    __syncthreads();
    for (int a = 0; a < n_angles; ++a) {
        const int buf = a & 1;
        ConstAngle &A = angles[a];
        const bool mx = A.m == 0;
        const int m0 = mx ? bx * BR : by * BR;
        const int c0 = mx ? by * BR : bx * BR;
        const int chunks = sh.chunks[buf];
        if (chunks == 0) {
            // no detector column of this angle meets the brick: only the next
            // angle's slice ranges are due (uniform branch over the block)
            if (t < BR && a + 1 < n_angles) bp_ranges(angles[a + 1], bx, by, cols, t, sh, buf ^ 1);
            __syncthreads();
            continue;
        }
        for (int c = 0; c < chunks; ++c) {
            // pass 1: thread (s = tx, jj = ty) splats detector column j onto kz
            // (a wave holds slots 4w .. 4w + 3 of all 16 slices: the empty slots,
            // the last ones of a slice, gather in the last waves, and a wave's
            // neighbouring slices read nearly the same detector columns)
            {
                const int s = tx, jj = ty;
                const int m = min(BR, sh.cnt[buf][s] - c * BR);  // valid columns of this chunk
                const bool valid = jj < m;
                const int pos = (!valid || sh.inc[buf][s]) ? jj : m - 1 - jj;
                const int j = sh.jlo[buf][s] + c * BR + jj;
                float *col = &sh.u[s * S::SLICE + pos * S::UCOL];
                if (valid) {
                    // (slots without a column are never read: q2 = 1e30)
                    auto zero_column = [&]() {
                        float2 *c2 = reinterpret_cast<float2 *>(col + UOFF);
#pragma unroll
                        for (int k = 0; k < BZ / 2; ++k) c2[k] = make_float2(0.0f, 0.0f);
                    };
                    // ray of detector row 0 in column j: slice n is met at the
                    // in-slice point (Az, Q2); row i adds i * Bz along z
                    const double n = (double)(m0 + s), jd = (double)j;
                    double Az, Bz, Q2;
                    if (A.cone) {
                        const double Dm = (A.P0[0] - A.S[0]) + jd * A.U[0];
                        const double Dz = (A.P0[1] - A.S[1]) + jd * A.U[1];
                        const double D2 = (A.P0[2] - A.S[2]) + jd * A.U[2];
                        const double tau = (n - A.S[0]) * rcp64(Dm);
                        Az = A.S[1] + tau * Dz;
                        Bz = tau * A.V[1];
                        Q2 = A.S[2] + tau * D2;
                    } else {
                        const double Pm = A.P0[0] + jd * A.U[0];
                        const double tau = (n - Pm) * rcp64(A.S[0]);
                        Az = (A.P0[1] + jd * A.U[1]) + tau * A.S[1];
                        Bz = A.V[1];
                        Q2 = (A.P0[2] + jd * A.U[2]) + tau * A.S[2];
                    }
                    sh.q2[s][pos] = (float)(Q2 - (double)c0);
                    // rows with t_i = Arel + i * Bz in (-1, BZ), bounded in FP32
                    // with a margin that exceeds its rounding: a row just outside
                    // is harmless (its t lies beyond the brick, see the splat
                    // below), and no row inside is lost
                    const double Arel = Az - (double)z0;
                    const float fB = (float)Bz;
                    const float fA = (float)Arel, iB = __builtin_amdgcn_rcpf(fB);
                    const float ra = (-1.0f - fA) * iB, rb = ((float)BZ - fA) * iB;
                    float lo = fminf(ra, rb), hi = fmaxf(ra, rb);
                    lo -= 1e-3f + 1e-5f * fabsf(lo);
                    hi += 1e-3f + 1e-5f * fabsf(hi);
                    const int r0 = (int)ceilf(fmaxf(lo, -1e9f)), r1 = (int)floorf(fminf(hi, 1e9f));
                    const int i0 = max(0, r0);
                    const int i1 = min(rows - 1, r1);
                    // no row of the range is cut off by the detector's edge: then
                    // rows less than two cells apart reach every cell of the brick
                    const bool covered = r0 >= 0 && r1 <= rows - 1 && r0 <= r1;
                    const float t0 = (float)(Arel + (double)i0 * Bz);
                    uint32_t off = (uint32_t)i0 * row_bytes + ((uint32_t)a * (uint32_t)cols + (uint32_t)j) * 4u;
                    float *u = col + UOFF;
                    // The rows are taken in increasing t (in the order of i if
                    // Bz > 0, reversed otherwise; the sign of Bz is that of the
                    // detector's v along z, the same for every column), in
                    // groups of four, the loads of the next group in flight
                    // while the current one is splatted, and no LDS cell is
                    // read. Rows past the range are read too (inside the
                    // buffer, or zero by the range check): their t lies beyond
                    // the brick (t >= BZ), and the clamp of kf to [-2, BZ]
                    // sends their weights to cells BZ, BZ + 1, which pass 2
                    // never reads. For t inside [-2, BZ + 1), kf = floor(t) and
                    // w = fract(t) are the exact bilinear split.
                    //
                    // Rows 1.01 .. 1.99 cells apart (|Bz|):
                    // a row's lower cell kf is at least the previous row's upper
                    // cell pk and at most one above it, and a cell takes at most
                    // two rows. The pending upper weight `carry` stays in a
                    // register, and each row writes cells (pk, pk + 1) once,
                    // final except for the pending one, with no LDS read:
                    //   kf == pk:     (carry + lo, hi)   hi is rewritten later
                    //   kf == pk + 1: (carry, lo)
                    // For a covered column every cell of the brick is written, so
                    // it needs no zeroing. kf is not clamped here: the rows of the
                    // range have kf in [-2, BZ], and the rows past it (kf above
                    // BZ) reach the LDS only through the write base min(pk, BZ),
                    // so they touch cells BZ, BZ + 1 only.
                    auto splat_carry = [&]() {
                        const bool up = fB > 0.0f;
                        const float dB = fabsf(fB);
                        const uint32_t step = up ? row_bytes : 0u - row_bytes;
                        float tb = up ? t0 : fmaf((float)(i1 - i0), fB, t0);
                        uint32_t o = up ? off : off + (uint32_t)(i1 - i0) * row_bytes;
                        int pk = min(max(floor_to_int(tb), -2), BZ);
                        float carry = 0.0f;
                        auto load4s = [&](uint32_t oo, float (&yv)[4]) {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                yv[r] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                                      rs, oo + (uint32_t)r * step, 0, 0));
                        };
                        auto carry4 = [&](float tg, const float (&yv)[4]) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const float ti = r == 0 ? tg : fmaf((float)r, dB, tg);
                                const int kf = floor_to_int(ti);
                                const float hiw = __builtin_amdgcn_fractf(ti) * yv[r];
                                const float low = yv[r] - hiw;
                                const bool eq = kf == pk;
                                const int b0 = min(pk, BZ);
                                u[b0] = carry + (eq ? low : 0.0f);
                                u[b0 + 1] = eq ? hiw : low;
                                pk = kf + 1;
                                carry = hiw;
                            }
                        };
                        float ya[4], yb[4];
                        const int nr = i1 - i0 + 1;
                        load4s(o, ya);
                        for (int i = 0; i < nr; i += 8) {
                            load4s(o + 4u * step, yb);
                            __builtin_amdgcn_sched_barrier(0);
                            carry4(tb, ya);
                            if (i + 4 >= nr) break;
                            load4s(o + 8u * step, ya);
                            __builtin_amdgcn_sched_barrier(0);
                            carry4(fmaf(4.0f, dB, tb), yb);
                            o += 8u * step;
                            tb = fmaf(8.0f, dB, tb);
                        }
                        u[min(pk, BZ)] = carry;
                    };
                    // Any row step: the two cells (p, p + 1) of the last row stay
                    // pending in registers (c0, c1); with d = kf - p a row adds
                    // to both (d = 0), finishes p (d = 1) or both (d >= 2), and
                    // writes (p, p + 1) once with their current values. Rows
                    // more than 1.99 cells apart leave cells between them that
                    // no row reaches: those columns are zeroed first, as are
                    // columns cut off by the detector's edge.
                    auto splat_carry2 = [&]() {
                        const bool up = fB > 0.0f;
                        const float dB = fabsf(fB);
                        const uint32_t step = up ? row_bytes : 0u - row_bytes;
                        float tb = up ? t0 : fmaf((float)(i1 - i0), fB, t0);
                        uint32_t o = up ? off : off + (uint32_t)(i1 - i0) * row_bytes;
                        int p = min(max(floor_to_int(tb), -2), BZ);
                        float c0 = 0.0f, c1 = 0.0f;
                        auto load4s = [&](uint32_t oo, float (&yv)[4]) {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                yv[r] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                                      rs, oo + (uint32_t)r * step, 0, 0));
                        };
                        auto carry4 = [&](float tg, const float (&yv)[4]) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const float ti = r == 0 ? tg : fmaf((float)r, dB, tg);
                                const int kf = min(max(floor_to_int(ti), -2), BZ);
                                const float hiw = __builtin_amdgcn_fractf(ti) * yv[r];
                                const float low = yv[r] - hiw;
                                const int d = kf - p;
                                const bool d0 = d == 0, d1 = d == 1;
                                const float v0 = d0 ? c0 + low : c0;
                                const float v1 = c1 + (d0 ? hiw : (d1 ? low : 0.0f));
                                u[p] = v0;
                                u[p + 1] = v1;
                                c0 = d0 ? v0 : (d1 ? v1 : low);
                                c1 = d0 ? v1 : hiw;
                                p = kf;
                            }
                        };
                        float ya[4], yb[4];
                        const int nr = i1 - i0 + 1;
                        load4s(o, ya);
                        for (int i = 0; i < nr; i += 8) {
                            load4s(o + 4u * step, yb);
                            __builtin_amdgcn_sched_barrier(0);
                            carry4(tb, ya);
                            if (i + 4 >= nr) break;
                            load4s(o + 8u * step, ya);
                            __builtin_amdgcn_sched_barrier(0);
                            carry4(fmaf(4.0f, dB, tb), yb);
                            o += 8u * step;
                            tb = fmaf(8.0f, dB, tb);
                        }
                        u[p] = c0;
                        u[p + 1] = c1;
                    };
                    if (__builtin_amdgcn_ballot_w64(fabsf(fB) > 1.01f && fabsf(fB) < 1.99f) ==
                        __builtin_amdgcn_read_exec()) {
                        if (!covered) zero_column();
                        splat_carry();
                    } else {
                        if (!covered || fabsf(fB) >= 1.99f) zero_column();
                        splat_carry2();
                    }
                } else {
                    sh.q2[s][pos] = 1e30f;
                }
            }
            // slice ranges of the next angle (the buffer was last read in the
            // previous angle, before its final barrier)
            if (c == 0 && t < BR && a + 1 < n_angles) bp_ranges(angles[a + 1], bx, by, cols, t, sh, buf ^ 1);
            __syncthreads();
            // pass 2: the stored columns with |Q_c2 - kc| < 1 follow the first
            // one with Q_c2 > kc - 1 (binary search over the increasing q2[s][.])
            {
                const int s = mx ? tx : ty;
                const float kc = (float)(mx ? ty : tx);
                const float *qs = sh.q2[s];
                int pos = 0;
#pragma unroll
                for (int step = BR / 2; step > 0; step >>= 1)
                    pos += (qs[pos + step - 1] <= kc - 1.0f) ? step : 0;
                pos += (qs[pos] <= kc - 1.0f) ? 1 : 0;  // all BR entries below
                const float *us = &sh.u[s * S::SLICE + UOFF];
                for (; pos < BR; ++pos) {
                    const float d = qs[pos] - kc;
                    if (!(d < 1.0f)) break;
                    const float w = 1.0f - fabsf(d);
                    const float2 *uc = reinterpret_cast<const float2 *>(us + pos * S::UCOL);
#pragma unroll
                    for (int k = 0; k < BZ / 2; ++k) {
                        const float2 uv = uc[k];
                        acc[2 * k] = fmaf(w, uv.x, acc[2 * k]);
                        acc[2 * k + 1] = fmaf(w, uv.y, acc[2 * k + 1]);
                    }
                }
            }
            __syncthreads();  // u and q2 are rewritten by the next chunk or angle
        }
    }
    if (x < v.nx && y < v.ny) {
        float *out = v.vol + (int64_t)b * v.vol_stride;
#pragma unroll
        for (int k = 0; k < BZ; ++k) {
            const int z = z0 + k;
            if (z < v.nz) emit(out, ((int64_t)z * v.ny + y) * v.nx + x, b, acc[k], v.ep);
        }
    }
}

template <int BZ>
__global__ void __launch_bounds__(BR *BR) __attribute__((amdgpu_waves_per_eu(BZ == 32 ? 4 : 1))) bp_separable(BpView v, int n_blocks, int shift)
{
    __shared__ __align__(16) BpShared<BZ> sh;
    const int logical = xcd_logical(shift);
    if (logical >= n_blocks) return;
    bp_separable_work<BZ>(v, logical, sh);
}

struct BpMulti {
    BpView v[MAX_FUSED];
    int begin[MAX_FUSED + 1];
    int n;
    int shift;  // chunk size of the XCD order (xcd_logical)
};

template <int BZ>
__global__ void __launch_bounds__(BR *BR) __attribute__((amdgpu_waves_per_eu(BZ == 32 ? 4 : 1))) bp_separable_multi(BpMulti m)
{
    __shared__ __align__(16) BpShared<BZ> sh;
    const int total = m.begin[m.n];
    const int logical = xcd_logical(m.shift);
    if (logical >= total) return;
    int k = 0;
    while (k + 1 < m.n && logical >= m.begin[k + 1]) ++k;
    bp_separable_work<BZ>(m.v[k], logical - m.begin[k], sh);
}

}  // namespace tsx
