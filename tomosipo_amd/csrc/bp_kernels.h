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
// bp_separable (kernels bp_wave, bp_wave_multi): for geometries whose
//   detector v axis is parallel to the in-slice axis c1 = z and whose
//   stepping axis is x or y (every circular, helical or jittered-source scan
//   with an upright detector). Then Q_c2 of a ray depends only on its
//   detector column j, and Q_z is affine in the row i, so per slice n
//     x[n, kz, kc2] = sum_j L(Q_c2(j; n) - kc2) * u[n, j, kz],
//     u[n, j, kz]   = sum_i L(A(j, n) + i B(j, n) - kz) * ly[i, j],
//   with ly = l * y precomputed. A block owns a brick of 16 x 16 columns by
//   BZ voxels along z; per angle it builds u for its 16 slices in LDS (pass
//   1: thread (slice, column) splats its detector column onto the z-grid,
//   rows in increasing z with the pending cells kept in registers, so that
//   every cell is written without being read back), stores the columns of a
//   slice in increasing order of Q_c2, and then every thread folds exactly
//   the columns with |Q_c2 - kc2| < 1 into the BZ voxels of its z-column
//   (pass 2; the first one found by binary search). Each wave owns four
//   slices in both passes, so the angle loop synchronises waves only with
//   themselves (see bp_wave_work). No atomics: the result is deterministic.
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
    // components of U and V (the plan's float copies)
    static __device__ __forceinline__ float u(const AngleParams &A, int k) { return A.fU[k]; }
    static __device__ __forceinline__ float v(const AngleParams &A, int k) { return A.fV[k]; }
};
template <>
struct GeneralMath<double> {
    static __device__ __forceinline__ double recip64(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double recip(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double root(double x) { return sqrt(x); }
    static __device__ __forceinline__ double u(const AngleParams &A, int k) { return A.U[k]; }
    static __device__ __forceinline__ double v(const AngleParams &A, int k) { return A.V[k]; }
};

// With PRESCALED (detector supersampling 1 only), proj holds the
// length-weighted projections l(a, i, j) * y[i, a, j] and the per-pixel ray
// length is not recomputed. CONE: the plan's beam (every angle of a plan is
// of the same kind), so that each kernel holds only its own set-up.
template <typename T, bool PRESCALED, bool CONE>
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
        T r1 = T(0), r2 = T(0), g = T(0);
        T A1j, A1i, A2j, A2i, fBj = T(0), fBi = T(0), bj, bi;
        bool unbounded = false;
        if constexpr (!CONE && sizeof(T) == sizeof(float)) {
            // parallel beam (float): the detector position is affine in the
            // voxel and the local map the same for every voxel of the angle
            if (!(A.par_b[0] >= 0.0f)) continue;  // singular map
            cj = fma(k2, A.par_c[3], fma(k1, A.par_c[2], fma(km, A.par_c[1], A.par_c[0])));
            ci = fma(k2, A.par_c[7], fma(k1, A.par_c[6], fma(km, A.par_c[5], A.par_c[4])));
            A1j = A.par_a[0];
            A1i = A.par_a[1];
            A2j = A.par_a[2];
            A2i = A.par_a[3];
            bj = A.par_b[0];
            bi = A.par_b[1];
            if constexpr (!PRESCALED) {
                r1 = (T)(A.S[1] * A.inv_S0);
                r2 = (T)(A.S[2] * A.inv_S0);
            }
        } else {
            T Tc, Bjc = T(0), Bic = T(0);
            if constexpr (CONE) {
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
                Tc = (T)(dn * A.inv_kappa);  // 1 / lam
                Bjc = M::u(A, 0) * Tc * iW0;
                Bic = M::v(A, 0) * Tc * iW0;
                g = iW0;
            } else {
                const double t0 = ((A.P0[0] - km) * A.N[0] + (A.P0[1] - k1) * A.N[1] +
                                   (A.P0[2] - k2) * A.N[2]) * A.inv_EN;
                const double R0 = km + t0 * A.S[0] - A.P0[0];
                const double R1 = k1 + t0 * A.S[1] - A.P0[1];
                const double R2 = k2 + t0 * A.S[2] - A.P0[2];
                cj = R0 * A.Bj[0] + R1 * A.Bj[1] + R2 * A.Bj[2];
                ci = R0 * A.Bi[0] + R1 * A.Bi[1] + R2 * A.Bi[2];
                r1 = (T)(A.S[1] * A.inv_S0);
                r2 = (T)(A.S[2] * A.inv_S0);
                Tc = T(1);
            }
            A1j = Tc * (M::u(A, 1) - r1 * M::u(A, 0));
            A1i = Tc * (M::v(A, 1) - r1 * M::v(A, 0));
            A2j = Tc * (M::u(A, 2) - r2 * M::u(A, 0));
            A2i = Tc * (M::v(A, 2) - r2 * M::v(A, 0));
            fBj = Bjc;
            fBi = Bic;
            // the local homography (offsets of O(1) pixels and voxels) in the
            // kernel's precision; only the voxel's own detector position
            // (cj, ci) needs FP64
            const T det = A1j * A2i - A1i * A2j;
            if (det == T(0)) continue;
            const T idet = GeneralMath<T>::recip(det);
            const T M00 = A2i * idet, M01 = -A1i * idet, M10 = -A2j * idet, M11 = A1j * idet;
            const T cv0 = fBj * M00 + fBi * M10, cv1 = fBj * M01 + fBi * M11;
            // The corners (s1, s2) = (+-1, +-1) of the square map to the
            // detector offsets (dj, di) = (M00 s1 + M01 s2, M10 s1 + M11 s2) /
            // (1 - cv0 s1 - cv1 s2); the image of the square is their convex
            // hull (the map is projective, with a denominator of one sign over
            // the square), so |dj| <= (|M00| + |M01|) / (1 - |cv0| - |cv1|) and
            // likewise |di| bound it: a box at most O(|cv|) wider than the
            // corners' own (cv is of the order of a voxel over the source
            // distance), from one reciprocal instead of four.
            const T den = T(1) - (fabs(cv0) + fabs(cv1));
            unbounded = !(den > T(1e-6));
            const T iden = GeneralMath<T>::recip(den);
            bj = (fabs(M00) + fabs(M01)) * iden;
            bi = (fabs(M10) + fabs(M11)) * iden;
        }
        int j0, j1, i0, i1;
        if (unbounded) {
            j0 = 0; j1 = cols - 1; i0 = 0; i1 = rows - 1;
        } else {
            const double mj = sub_margin + 1e-3 + (double)bj, mi = sub_margin + 1e-3 + (double)bi;
            j0 = (int)fmax(0.0, ceil(cj - mj));
            j1 = (int)fmin((double)(cols - 1), floor(cj + mj));
            i0 = (int)fmax(0.0, ceil(ci - mi));
            i1 = (int)fmin((double)(rows - 1), floor(ci + mi));
        }
        if (j0 > j1 || i0 > i1) continue;
        const T dj0 = (T)((double)j0 - cj), di0 = (T)((double)i0 - ci);
        if constexpr (PRESCALED) {
            // detector supersampling 1 (the only case with length-weighted
            // projections): one ray per pixel, through its centre
            T acc_a = T(0);
            for (int i = i0; i <= i1; ++i) {
                const T *row = pr + ((int64_t)i * n_angles + a) * cols;
                const T di = di0 + (T)(i - i0);
                const T e = di * fBi + T(1), b1 = di * A1i, b2 = di * A2i;
                for (int j = j0; j <= j1; ++j) {
                    const T dj = dj0 + (T)(j - j0);
                    const T iden = GeneralMath<T>::recip(dj * fBj + e);
                    const T dq1 = (dj * A1j + b1) * iden;
                    const T dq2 = (dj * A2j + b2) * iden;
                    const T w1 = T(1) - fabs(dq1), w2 = T(1) - fabs(dq2);
                    const T w = (w1 > T(0) ? w1 : T(0)) * (w2 > T(0) ? w2 : T(0));
                    acc_a += w * row[j];
                }
            }
            acc += acc_a;
            continue;
        }
        const T fr1 = r1, fr2 = r2, fg = g;
        const T h0 = (T)(A.h[0] * A.h[0]), h1 = (T)(A.h[1] * A.h[1]), h2 = (T)(A.h[2] * A.h[2]);
        const T len_par = GeneralMath<T>::root(h0 + h1 * fr1 * fr1 + h2 * fr2 * fr2);
        const T tinv = (T)1 / (T)ss;
        T acc_a = T(0);
        for (int i = i0; i <= i1; ++i) {
            const T *row = pr + ((int64_t)i * n_angles + a) * cols;
            for (int j = j0; j <= j1; ++j) {
                const T yv = row[j];
                T wsum = T(0);
                for (int p = 0; p < ss; ++p) {
                    const T di = di0 + (T)(i - i0) + (((T)p + T(0.5)) * tinv - T(0.5));
                    for (int q = 0; q < ss; ++q) {
                        const T dj = dj0 + (T)(j - j0) + (((T)q + T(0.5)) * tinv - T(0.5));
                        const T den = dj * fBj + (di * fBi + T(1));
                        const T iden = GeneralMath<T>::recip(den);
                        const T dq1 = (dj * A1j + di * A1i) * iden;
                        const T dq2 = (dj * A2j + di * A2i) * iden;
                        const T w1 = T(1) - fabs(dq1), w2 = T(1) - fabs(dq2);
                        const T w = (w1 > T(0) ? w1 : T(0)) * (w2 > T(0) ? w2 : T(0));
                        T len = len_par;
                        if constexpr (CONE) {
                            const T s1 = dq1 * fg + fr1, s2 = dq2 * fg + fr2;
                            len = GeneralMath<T>::root(h0 + h1 * s1 * s1 + h2 * s2 * s2);
                        }
                        wsum += w * len;
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
// World length per slice step of the ray through pixel (i, j) of angle A.
__device__ __forceinline__ float ray_length(const AngleParams &A, int i, int j)
{
    float D[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float P = (float)(A.P0[c] + (double)j * A.U[c] + (double)i * A.V[c]);
        D[c] = A.cone ? P - (float)A.S[c] : (float)A.S[c];
    }
    const float inv = 1.0f / D[0];
    const float a1 = D[1] * inv, a2 = D[2] * inv;
    return (D[0] == 0.0f) ? 0.0f
        : sqrtf((float)(A.h[0] * A.h[0]) + (float)(A.h[1] * A.h[1]) * a1 * a1 +
                (float)(A.h[2] * A.h[2]) * a2 * a2);
}

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
        const float len = ray_length(angles[a], i, j);
        const int64_t o = ia * cols + j;
        out[b * out_stride + o] = len * proj[b * proj_stride + o];
    }
}

// Length-weighted projections for the separable backprojector, transposed so
// that the rows of one detector column are contiguous, for the kn angles
// sorted[0 .. kn - 1] (each with its index in the projections in `pad`):
//   out[b][k][j][RPAD + i] = l(a, i, j) * proj[b][i][a][j],  a = sorted[k].pad,
// each column padded with RPAD zero rows on both sides (rows -RPAD .. -1 and
// rows .. rows + RPAD - 1), so that loads of four consecutive rows starting
// anywhere in [-RPAD, rows - 1] stay inside the column. Block (x, y): tile of
// 32 rows x 32 columns (x = row tile + tiles_i * column tile) of the lines y,
// y + gridDim.y, ... of the kn * batch (entry, angle) pairs, through LDS.
constexpr int RPAD = 3;

// One (tile, angle) pair of scale_by_length_t: the tile (i0, j0) of sorted
// angle k of one batch entry (proj and out moved to the entry). up > 1: the
// plan's detector is a detector of (rows / up) x (cols / up) pixels refined
// up times, and proj holds that coarse detector's projections; fine pixel
// (i, j) takes coarse pixel (i / up, j / up) with weight 1 / up^2.
__device__ __forceinline__ void scale_t_tile(const AngleParams *__restrict__ sorted, int k, int n_angles,
                                             int rows, int cols, const float *__restrict__ proj,
                                             float *__restrict__ out, int i0, int j0, float (&tile)[32][33],
                                             int up = 1)
{
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    const int rp = rows + 2 * RPAD;
    const AngleParams &A = sorted[k];
    const int a = A.pad;
    __syncthreads();  // the tile of the block's previous pair has been written out
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        float v = 0.0f;
        if (i < rows && j < cols) {
            const float y = up == 1 ? proj[((int64_t)i * n_angles + a) * cols + j]
                                    : proj[((int64_t)(i / up) * n_angles + a) * (cols / up) + j / up] /
                                          (float)(up * up);
            v = ray_length(A, i, j) * y;
        }
        tile[r][tx] = v;
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
        const int j = j0 + c;
        if (j >= cols) continue;
        float *col = out + ((int64_t)k * cols + j) * rp;
        if (i0 + tx < rows) col[RPAD + i0 + tx] = tile[tx][c];
        if (tx < RPAD) {
            if (i0 == 0) col[tx] = 0.0f;
            if (i0 + 32 >= rows) col[RPAD + rows + tx] = 0.0f;
        }
    }
}

__global__ void __launch_bounds__(256)
scale_by_length_t(const AngleParams *__restrict__ sorted, int kn, int n_angles, int rows, int cols,
                  const float *__restrict__ proj, int64_t proj_stride, float *__restrict__ out,
                  int64_t out_stride, int64_t lines, int up)
{
    __shared__ float tile[32][33];
    const int tiles_i = (rows + 31) / 32;
    const int i0 = (int)(blockIdx.x % tiles_i) * 32, j0 = (int)(blockIdx.x / tiles_i) * 32;
    for (int64_t line = blockIdx.y; line < lines; line += gridDim.y) {
        const int64_t b = line / kn;
        const int k = (int)(line - b * kn);
        scale_t_tile(sorted, k, n_angles, rows, cols, proj + b * proj_stride, out + b * out_stride, i0, j0, tile,
                     up);
    }
}

// scale_by_length_t of several plans' projections in one launch: job k is
// row y = k of the grid, whose blocks x < blocks[k] each take one (tile,
// angle) pair (the rest of the row returns at once; no table lookup in
// these short blocks).
struct ScaleJob {
    const AngleParams *sorted;
    int kn, n_angles, rows, cols;
    const float *proj;
    float *out;
    int blocks;
};

struct ScaleJobs {
    ScaleJob job[MAX_FUSED];
    int n;
};

__host__ __device__ inline int64_t scale_job_blocks(const ScaleJob &j)
{
    return (int64_t)((j.rows + 31) / 32) * ((j.cols + 31) / 32) * j.kn;
}

__global__ void __launch_bounds__(256) scale_by_length_t_multi(ScaleJobs S)
{
    __shared__ float tile[32][33];
    const ScaleJob &J = S.job[blockIdx.y];
    const int r = blockIdx.x;
    if (r >= J.blocks) return;
    const int tiles_i = (J.rows + 31) / 32, tiles = tiles_i * ((J.cols + 31) / 32);
    const int t = r % tiles;
    scale_t_tile(J.sorted, r / tiles, J.n_angles, J.rows, J.cols, J.proj, J.out, (t % tiles_i) * 32,
                 (t / tiles_i) * 32, tile);
}

// A volume in the layout with axis swap - 1 (x or y) and z exchanged, tmp
// ([x][y][z] for swap 1, [y][z][x] for swap 2; entries ts floats apart),
// written to out ([z][y][x] of extents nx, ny, nz; entries vs apart) through
// the epilogue. The exchange is its own inverse: with the extents of the
// swapped layout it also moves a volume into that layout. swap 1: a block per
// 32 x 32 tile of the (z, x) plane of one y, through LDS (block id = tile +
// tiles * y); swap 2: a block per 256 x-elements of one (z, y) row, whose
// x-rows are contiguous in both. Block ids start at `first` (launches over
// ranges of them stay within the dispatch limit); blockIdx.z: batch entry.
__global__ void __launch_bounds__(256) unswap_volume(const float *__restrict__ tmp, int64_t ts,
                                                      float *__restrict__ out, int64_t vs, int nx, int ny,
                                                      int nz, int swap, int64_t first, Epilogue ep)
{
    const int64_t b = blockIdx.z;
    const float *src = tmp + b * ts;
    float *dst = out + b * vs;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    const int64_t id = first + blockIdx.x;
    if (swap == 1) {
        __shared__ float tile[32][33];
        const int tiles_x = (nx + 31) / 32, tiles = tiles_x * ((nz + 31) / 32);
        const int y = (int)(id / tiles), t = (int)(id % tiles);
        const int x0 = (t % tiles_x) * 32, z0 = (t / tiles_x) * 32;
        for (int r = ty; r < 32; r += 8) {  // tmp rows x0 + r, z contiguous
            const int x = x0 + r, z = z0 + tx;
            tile[r][tx] = (x < nx && z < nz) ? src[((int64_t)x * ny + y) * nz + z] : 0.0f;
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {  // out rows z0 + r, x contiguous
            const int z = z0 + r, x = x0 + tx;
            if (x < nx && z < nz) emit(dst, ((int64_t)z * ny + y) * nx + x, b, tile[tx][r], ep);
        }
    } else {
        const int xb = (nx + 255) / 256;
        const int64_t row = id / xb;  // (z, y) of out
        const int x = (int)(id % xb) * 256 + ty * 32 + tx;
        if (x >= nx) return;
        const int z = (int)(row / ny), y = (int)(row % ny);
        emit(dst, row * nx + x, b, src[((int64_t)y * nz + z) * nx + x], ep);
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
    const AngleParams *sorted;  // the angles ordered by stepping axis, original index in `pad`
    int n_angles, rows, cols, nx, ny, nz, bricks_x, bricks_y, bricks_z, batch;
    const uint32_t *rt;  // slice-range table (bp_range_table), rt_stride entries per angle
    int64_t rt_stride;
    float *vol;
    int64_t vol_stride;
    const float *ly;    // length-weighted projections, batch entries ly_bytes apart
    uint32_t ly_bytes;
    Epilogue ep;
    // angle parts: the sums over the angles are taken in parts of part_len
    // consecutive angles, added in order. partial != nullptr: a block per
    // (brick, entry, part) writes its part's sums to partial (entry e, part s
    // at (e * parts + s) * nx * ny * nz) for bp_parts_reduce; otherwise one
    // block per (brick, entry) runs the parts in turn (kernels with PARTS).
    int parts, part_len;
    float *partial;
};

// LDS layout of the separable backprojector: columns are UCOL = BZ + 2
// floats apart and slices SLICE floats apart (SLICE = 36 mod 64); the cell kz
// of the column in slot pos of slice s is at s * SLICE + pos * UCOL + UOFF +
// kz. Only the cells kz = 0 .. BZ - 1 are read. The splat clamps the lower
// cell of a row to [-2, BZ], so the weights of rows outside the brick land in
// cells -2, -1 of their column or in cells BZ, BZ + 1; a column's cells
// BZ - 2, BZ - 1 share the next column's (unused) cells -4, -3, and its cells
// BZ, BZ + 1 the next column's cells -2, -1, so no real cell is written by
// another column. The strides were chosen with a simulation of the LDS
// addresses of an earlier, block-synchronous form of the kernel on a
// cone-beam geometry (64 dword banks; 7 % fewer bank cycles than columns
// BZ + 4 apart) and kept; every column stays 8-byte aligned so that pass 2
// reads its cells as float2.
constexpr int UOFF = 4;  // kz = 0 sits at offset 4 of a column

template <int BZ>
struct BpShared {
    static constexpr int UCOL = BZ + 2;
    static constexpr int SLICE = (BR * UCOL + 27) / 64 * 64 + 36;  // first stride >= BR * UCOL + 2, = 36 mod 64
    static_assert(UCOL % 2 == 0 && SLICE % 64 == 36 && SLICE >= BR * UCOL + 2, "LDS strides");
    static constexpr int TROW = BR + 1;      // staging of the accumulator transpose:
    static constexpr int TCELL = BR * TROW;  // [kz][row][col], rows BR + 1 apart (in u)
    static_assert(BZ * TCELL <= BR * SLICE, "the transpose staging fits in u");
    float u[BR * SLICE];
    float q2[BR][20];  // Q_c2 - c0 per stored column, increasing; 1e30 if unused
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

// Slice-range table of the separable backprojector: for every angle k of the
// axis-ordered table, every column-block cb of 16 voxels along the in-slice
// axis c2 and every slice n along the stepping axis (rounded up to whole
// bricks), entry k * stride + cb * nm + n (nm = slices rounded up to 16)
// holds the detector columns that meet the brick's slice n, as
// jlo | cnt << 16 | inc << 31 (first column, count, 1 if Q_c2 increases with
// the column; needs fewer than 32768 detector columns). Depends only on the
// geometry: built once per plan, so the kernel reads one word per lane and
// angle instead of evaluating the FP64 column map twice.
constexpr int BP_RT_MAX_COLS = 32767;

// Entries per angle: bricks along x times bricks along y times BR slices
// (the same for x- and y-stepping angles).
__host__ __device__ inline int64_t bp_rt_stride(int nx, int ny)
{
    return (int64_t)((nx + BR - 1) / BR) * ((ny + BR - 1) / BR) * BR;
}

__global__ void bp_range_table(const AngleParams *__restrict__ sorted, int cols, int nx, int ny, int64_t stride,
                               uint32_t *__restrict__ out)
{
    const int k = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= stride) return;
    const AngleParams &A = sorted[k];
    const bool mx = A.m == 0;
    const int nm = mx ? (nx + BR - 1) / BR * BR : (ny + BR - 1) / BR * BR;
    const int ncb = mx ? (ny + BR - 1) / BR : (nx + BR - 1) / BR;
    if (e >= (int64_t)nm * ncb) return;
    const int cb = (int)(e / nm), n = (int)(e % nm);
    const int bx = mx ? n / BR : cb, by = mx ? cb : n / BR;
    int jlo, cnt, inc;
    bp_slice_range(A, bx, by, cols, n % BR, jlo, cnt, inc);
    out[(int64_t)k * stride + e] = (uint32_t)jlo | ((uint32_t)cnt << 16) | ((uint32_t)inc << 31);
}

// Rows ig .. ig + 3 of a column of the transposed length-weighted
// projections (byte offset colbase + 4 ig), in increasing t: in the order of
// the rows when UP, reversed otherwise.
template <bool UP>
__device__ __forceinline__ void bp_rows4(__amdgpu_buffer_rsrc_t rs, uint32_t off, float (&y)[4])
{
    const float4 v = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
    y[0] = UP ? v.x : v.w;
    y[1] = UP ? v.y : v.z;
    y[2] = UP ? v.z : v.y;
    y[3] = UP ? v.w : v.x;
}

// Splat of the rows i0 .. i1 of one column onto the cells u[kz] (the bilinear
// transpose along z), rows 1.01 .. 1.99 cells apart (|Bz|): a row's lower
// cell kf is at least the previous row's upper cell pk and at most one above
// it, and a cell takes at most two rows. The pending upper weight `carry`
// stays in a register, and each row writes cells (pk, pk + 1) once, final
// except for the pending one, with no LDS read:
//   kf == pk:     (carry + lo, hi)   hi is rewritten later
//   kf == pk + 1: (carry, lo)
// For t inside [-2, BZ + 1), kf = floor(t) and w = fract(t) are the exact
// bilinear split. The rows of the range have t in (-1 - e, BZ + e) (e the
// small margin of the row range), so kf is in [-2, BZ] and only the last row
// of the range can have kf = BZ: every row of a full group of four writes
// cells up to BZ + 1 without a clamp. The last, partial group also takes up
// to three rows past the range (their t lies beyond the brick, t >= BZ); it
// writes through the base min(pk, BZ), which sends those rows to cells BZ,
// BZ + 1 (the next column's unused cells -2, -1), never read by pass 2.
// Group g starts at row i0 + 4g (UP) or i1 - 3 - 4g, never below i0 - 3: in
// the column's zero padding at worst. The load of the group after the last
// may fall beyond the column (its values are not used).
template <int BZ, bool UP, bool CLAMP>
__device__ __forceinline__ void bp_carry4(float tg, float dB, const float (&y)[4], float *u, int &pk, float &carry)
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float ti = r == 0 ? tg : fmaf((float)r, dB, tg);
        const int kf = floor_to_int(ti);
        const float hiw = __builtin_amdgcn_fractf(ti) * y[r];
        const float low = y[r] - hiw;
        const bool eq = kf == pk;
        const int b0 = CLAMP ? min(pk, BZ) : pk;
        u[b0] = carry + (eq ? low : 0.0f);
        u[b0 + 1] = eq ? hiw : low;
        pk = kf + 1;
        carry = hiw;
    }
}

template <int BZ, bool UP>
__device__ __forceinline__ void bp_splat_carry(__amdgpu_buffer_rsrc_t rs, uint32_t colbase, int i0, int i1, float t0,
                                               float fB, float *u)
{
    const float dB = fabsf(fB);
    const int nr = i1 - i0 + 1;
    float tb = UP ? t0 : fmaf((float)(i1 - i0), fB, t0);  // t of the first row taken
    int pk = min(max(floor_to_int(tb), -2), BZ);
    float carry = 0.0f;
    if (nr > 0) {
        // two groups per iteration, loaded alternately into ya and yb, each
        // load in flight while the other group is splatted
        const uint32_t step = UP ? 16u : 0u - 16u;
        uint32_t o = colbase + (uint32_t)(UP ? i0 : i1 - 3) * 4u;
        float ya[4], yb[4];
        bp_rows4<UP>(rs, o, ya);
        const int nfull = nr >> 2;
        int g = 0;
        for (; g + 1 < nfull; g += 2) {
            bp_rows4<UP>(rs, o + step, yb);
            __builtin_amdgcn_sched_barrier(0);
            bp_carry4<BZ, UP, false>(tb, dB, ya, u, pk, carry);
            bp_rows4<UP>(rs, o + 2u * step, ya);
            __builtin_amdgcn_sched_barrier(0);
            bp_carry4<BZ, UP, false>(fmaf(4.0f, dB, tb), dB, yb, u, pk, carry);
            o += 2u * step;
            tb = fmaf(8.0f, dB, tb);
        }
        // ya holds group g: the last full group (nfull odd) or the partial one
        if (g < nfull) {
            bp_rows4<UP>(rs, o + step, yb);
            __builtin_amdgcn_sched_barrier(0);
            bp_carry4<BZ, UP, false>(tb, dB, ya, u, pk, carry);
            if (nr & 3) bp_carry4<BZ, UP, true>(fmaf(4.0f, dB, tb), dB, yb, u, pk, carry);
        } else if (nr & 3) {
            bp_carry4<BZ, UP, true>(tb, dB, ya, u, pk, carry);
        }
    }
    u[min(pk, BZ)] = carry;
}

// Any row step: the two cells (p, p + 1) of the last row stay pending in
// registers (c0, c1); with d = kf - p a row adds to both (d = 0), finishes p
// (d = 1) or both (d >= 2), and writes (p, p + 1) once with their current
// values; kf is clamped to [-2, BZ] for every row. Rows more than 1.99 cells
// apart leave cells between them that no row reaches: the caller zeroes those
// columns first, as it does columns cut off by the detector's edge.
template <int BZ, bool UP>
__device__ __forceinline__ void bp_splat_carry2(__amdgpu_buffer_rsrc_t rs, uint32_t colbase, int i0, int i1, float t0,
                                                float fB, float *u)
{
    const float dB = fabsf(fB);
    const int nr = i1 - i0 + 1;
    float tb = UP ? t0 : fmaf((float)(i1 - i0), fB, t0);
    int p = min(max(floor_to_int(tb), -2), BZ);
    float c0 = 0.0f, c1 = 0.0f;
    if (nr > 0) {
        const uint32_t step = UP ? 16u : 0u - 16u;
        uint32_t o = colbase + (uint32_t)(UP ? i0 : i1 - 3) * 4u;
        float cur[4], nxt[4];
        bp_rows4<UP>(rs, o, cur);
        for (int i = 0; i < nr; i += 4) {
            o += step;
            bp_rows4<UP>(rs, o, nxt);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float ti = r == 0 ? tb : fmaf((float)r, dB, tb);
                const int kf = min(max(floor_to_int(ti), -2), BZ);
                const float hiw = __builtin_amdgcn_fractf(ti) * cur[r];
                const float low = cur[r] - hiw;
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
#pragma unroll
            for (int r = 0; r < 4; ++r) cur[r] = nxt[r];
            tb = fmaf(4.0f, dB, tb);
        }
    }
    u[p] = c0;
    u[p + 1] = c1;
}

// Pass 1 of the separable backprojector for one detector column j of one
// slice n_slice: the bilinear transpose along z of the column's rows onto the
// BZ cells of the brick (starting at z0), written to col[UOFF + kz], and the
// column's in-slice position Q_c2 - c0 to q2. `a` is the angle's index in the
// projections; rs addresses their transposed, padded, length-weighted copy
// (scale_by_length_t).
template <int BZ, typename AP>
__device__ __forceinline__ void bp_column_splat(const AP &A, __amdgpu_buffer_rsrc_t rs, int a, int cols, int rows,
                                                int j, int n_slice, int c0, int z0, float *col, float &q2)
{
    // (slots without a column are never read: q2 = 1e30)
    auto zero_column = [&]() {
        float2 *c2 = reinterpret_cast<float2 *>(col + UOFF);
#pragma unroll
        for (int k = 0; k < BZ / 2; ++k) c2[k] = make_float2(0.0f, 0.0f);
    };
    // ray of detector row 0 in column j: slice n is met at the
    // in-slice point (Az, Q2); row i adds i * Bz along z
    const double n = (double)n_slice, jd = (double)j;
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
    q2 = (float)(Q2 - (double)c0);
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
    float *u = col + UOFF;
    // The rows are taken in increasing t (in the order of i if Bz > 0,
    // reversed otherwise; the sign of Bz is that of the detector's v along z,
    // the same for every column: UP is uniform), in groups of four
    // consecutive rows read as one 16-byte load from the column's rows in the
    // transposed, padded layout of scale_by_length_t, the next group in
    // flight while the current one is splatted, and no LDS cell is read.
    const uint32_t colbase =
        (((uint32_t)a * (uint32_t)cols + (uint32_t)j) * (uint32_t)(rows + 2 * RPAD) + (uint32_t)RPAD) * 4u;
    if (__builtin_amdgcn_ballot_w64(fabsf(fB) > 1.01f && fabsf(fB) < 1.99f) == __builtin_amdgcn_read_exec()) {
        if (!covered) zero_column();
        if (A.V[1] > 0.0)
            bp_splat_carry<BZ, true>(rs, colbase, i0, i1, t0, fB, u);
        else
            bp_splat_carry<BZ, false>(rs, colbase, i0, i1, t0, fB, u);
    } else {
        if (!covered || fabsf(fB) >= 1.99f) zero_column();
        if (A.V[1] > 0.0)
            bp_splat_carry2<BZ, true>(rs, colbase, i0, i1, t0, fB, u);
        else
            bp_splat_carry2<BZ, false>(rs, colbase, i0, i1, t0, fB, u);
    }
}

// ---------------------------------------------------------------------------
// The separable backprojector with wave-local slices. Thread (tx, ty) of the
// 16 x 16 block works on slice s = ty of the brick in both passes (pass 1:
// detector column slot tx; pass 2: in-slice voxel column kc = tx), so each
// wave (slices 4w .. 4w + 3) exchanges LDS data only with itself and the
// angle loop has no block barrier: a wave finishes its pass 2 of one angle
// and starts the next while the others are still at work. The voxel column a
// thread accumulates is (x, y) = (c0 + tx, m0 + ty) for y-stepping angles and
// (m0 + ty, c0 + tx) for x-stepping ones, so the angles are taken from the
// plan's table ordered by stepping axis (x first) and the accumulators are
// moved between threads (tx, ty) and (ty, tx) through LDS where the axis
// changes (once per launch) and at the end if the last angle steps along x.
// Each wave computes the slice ranges of its own four slices for the next
// angle (in registers) and runs as many column chunks as those slices need.

// Orders the LDS accesses of one wave around this point. LDS operations of a
// wave complete in issue order, so the fences (wavefront scope) only keep the
// compiler from moving LDS reads and writes across it.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc of thread (tx, ty) <- acc of thread (ty, tx), through LDS (block-wide;
// called by every thread of the block at the same point of the angle loop).
template <int BZ>
__device__ __forceinline__ void bp_wave_transpose(float (&acc)[BZ], float *stage, int tx, int ty)
{
    using S = BpShared<BZ>;
    __syncthreads();  // every wave is done with u
#pragma unroll
    for (int k = 0; k < BZ; ++k) stage[k * S::TCELL + ty * S::TROW + tx] = acc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BZ; ++k) acc[k] = stage[k * S::TCELL + tx * S::TROW + ty];
    __syncthreads();  // u is written again by pass 1
}

// The lane's entry of the slice-range table for angle k (its slice ty of the
// block's brick; the c2 block and slice index follow the angle's axis).
__device__ __forceinline__ uint32_t bp_wave_entry(const BpView &v, int k, bool mx, int bx, int by, int ty)
{
    const int64_t e = mx ? (int64_t)by * (v.bricks_x * BR) + bx * BR + ty
                         : (int64_t)bx * (v.bricks_y * BR) + by * BR + ty;
    return v.rt[(int64_t)k * v.rt_stride + e];
}

// Decodes a table entry; chunks = the wave's number of column chunks (the
// largest over its four slices, lanes 16 and 32 apart).
__device__ __forceinline__ void bp_wave_decode(uint32_t w, int &jlo, int &cnt, int &inc, int &chunks)
{
    jlo = (int)(w & 0xffffu);
    cnt = (int)((w >> 16) & 0x7fffu);
    inc = (int)(w >> 31);
    int ch = (cnt + BR - 1) / BR;
    ch = max(ch, __shfl_xor(ch, 16));
    ch = max(ch, __shfl_xor(ch, 32));
    chunks = __builtin_amdgcn_readfirstlane(ch);
}

template <int BZ, bool PARTS>
__device__ __forceinline__ void bp_wave_work(const BpView &v, int logical, BpShared<BZ> &sh)
{
    using S = BpShared<BZ>;
    const int n_angles = v.n_angles, rows = v.rows, cols = v.cols;
    const int bx = logical % v.bricks_x;
    const int by = (logical / v.bricks_x) % v.bricks_y;
    const int bz = (logical / (v.bricks_x * v.bricks_y)) % v.bricks_z;
    int b = logical / (v.bricks_x * v.bricks_y * v.bricks_z);
    int ka = 0, kb = n_angles, part = 0;  // this block's angles
    if (v.partial) {
        part = b % v.parts;
        b /= v.parts;
        ka = part * v.part_len;
        kb = min(n_angles, ka + v.part_len);
    }
    if (b >= v.batch) return;
    // angle parameters through a constant-address-space pointer: the uniform
    // ones become scalar loads
    typedef const __attribute__((address_space(4))) AngleParams ConstAngle;
    ConstAngle *__restrict__ angles = (ConstAngle *)v.sorted;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int z0 = bz * BZ;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)uniform_ptr(v.ly + (int64_t)b * ((int64_t)v.ly_bytes / 4)), (short)0,
        (int)__builtin_amdgcn_readfirstlane(v.ly_bytes), 0x00020000);
    float acc[BZ];
    float tot[PARTS ? BZ : 1];  // sums of the finished parts
#pragma unroll
    for (int k = 0; k < BZ; ++k) acc[k] = 0.0f;
    if (PARTS) {
#pragma unroll
        for (int k = 0; k < BZ; ++k) tot[PARTS ? k : 0] = 0.0f;
    }
    int mode = angles[ka].m;  // stepping axis whose voxel mapping acc follows
    // slice ranges from the table, read one angle ahead
    uint32_t next = bp_wave_entry(v, ka, mode == 0, bx, by, ty);
    for (int k = ka; k < kb; ++k) {
        ConstAngle &A = angles[k];
        const int am = A.m;
        if (PARTS && k > ka && k % v.part_len == 0) {
            // a part ends: its sums (in the final voxel mapping) join the total
            if (mode == 0) bp_wave_transpose<BZ>(acc, sh.u, tx, ty);
#pragma unroll
            for (int q = 0; q < BZ; ++q) {
                tot[PARTS ? q : 0] += acc[q];
                acc[q] = 0.0f;
            }
            mode = am;
        }
        if (am != mode) {
            bp_wave_transpose<BZ>(acc, sh.u, tx, ty);
            mode = am;
        }
        int cur_jlo, cur_cnt, cur_inc, cur_chunks;
        bp_wave_decode(next, cur_jlo, cur_cnt, cur_inc, cur_chunks);
        if (k + 1 < kb) next = bp_wave_entry(v, k + 1, angles[k + 1].m == 0, bx, by, ty);
        const bool mx = am == 0;
        const int m0 = mx ? bx * BR : by * BR;
        const int c0 = mx ? by * BR : bx * BR;
        const int a = k;  // the angle's transposed projections in ly
        for (int c = 0; c < cur_chunks; ++c) {
            // pass 1: thread (slice ty, slot tx) splats detector column j onto kz
            {
                const int m = min(BR, cur_cnt - c * BR);  // valid columns of this chunk
                const bool valid = tx < m;
                const int pos = (!valid || cur_inc) ? tx : m - 1 - tx;
                const int j = cur_jlo + c * BR + tx;
                float *col = &sh.u[ty * S::SLICE + pos * S::UCOL];
                if (valid)
                    bp_column_splat<BZ>(A, rs, a, cols, rows, j, m0 + ty, c0, z0, col, sh.q2[ty][pos]);
                else
                    sh.q2[ty][pos] = 1e30f;  // (slots without a column are never read)
            }
            wave_lds_sync();
            // pass 2: the stored columns with |Q_c2 - kc| < 1 follow the first
            // one with Q_c2 > kc - 1 (binary search over the increasing q2[ty][.])
            {
                const float kc = (float)tx;
                const float *qs = sh.q2[ty];
                int pos = 0;
#pragma unroll
                for (int step = BR / 2; step > 0; step >>= 1)
                    pos += (qs[pos + step - 1] <= kc - 1.0f) ? step : 0;
                pos += (qs[pos] <= kc - 1.0f) ? 1 : 0;  // all BR entries below
                const float *us = &sh.u[ty * S::SLICE + UOFF];
                for (; pos < BR; ++pos) {
                    const float d = qs[pos] - kc;
                    if (!(d < 1.0f)) break;
                    const float w = 1.0f - fabsf(d);
                    const float2 *uc = reinterpret_cast<const float2 *>(us + pos * S::UCOL);
#pragma unroll
                    for (int q = 0; q < BZ / 2; ++q) {
                        const float2 uv = uc[q];
                        acc[2 * q] = fmaf(w, uv.x, acc[2 * q]);
                        acc[2 * q + 1] = fmaf(w, uv.y, acc[2 * q + 1]);
                    }
                }
            }
            wave_lds_sync();  // u and q2 are rewritten by the next chunk or angle
        }
    }
    if (mode == 0) bp_wave_transpose<BZ>(acc, sh.u, tx, ty);
    if (PARTS) {
#pragma unroll
        for (int q = 0; q < BZ; ++q) acc[q] = tot[PARTS ? q : 0] + acc[q];
    }
    // thread (tx, ty) now holds voxel column (x, y) = (bx * BR + tx, by * BR + ty)
    const int x = bx * BR + tx, y = by * BR + ty;
    if (x < v.nx && y < v.ny) {
        if (v.partial) {
            const int64_t ve = (int64_t)v.nx * v.ny * v.nz;
            float *out = v.partial + ((int64_t)b * v.parts + part) * ve;
#pragma unroll
            for (int q = 0; q < BZ; ++q) {
                const int z = z0 + q;
                if (z < v.nz) out[((int64_t)z * v.ny + y) * v.nx + x] = acc[q];
            }
        } else {
            float *out = v.vol + (int64_t)b * v.vol_stride;
#pragma unroll
            for (int q = 0; q < BZ; ++q) {
                const int z = z0 + q;
                if (z < v.nz) emit(out, ((int64_t)z * v.ny + y) * v.nx + x, b, acc[q], v.ep);
            }
        }
    }
}

// The sums of the angle parts of bp_wave (partial, parts per entry) added in
// order and written through the epilogue: grid (voxels / 256, entries).
__global__ void __launch_bounds__(256)
bp_parts_reduce(const float *__restrict__ partial, int parts, int64_t ve, float *vol, int64_t vs, Epilogue ep)
{
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (o >= ve) return;
    const float *pp = partial + (int64_t)b * parts * ve + o;
    float t = pp[0];
    for (int s = 1; s < parts; ++s) t += pp[(int64_t)s * ve];
    emit(vol + (int64_t)b * vs, o, b, t, ep);
}

// Waves per SIMD the register allocation is held to: 4 (16 per CU, four
// blocks) for 32-deep bricks, 5 for 24-deep ones (their 28 KB of LDS fit five
// blocks per CU), no bound for 16-deep ones.
constexpr int bp_waves_per_eu(int bz) { return bz == 32 ? 4 : bz == 24 ? 5 : 1; }

template <int BZ>
__global__ void __launch_bounds__(BR *BR) __attribute__((amdgpu_waves_per_eu(bp_waves_per_eu(BZ)))) bp_wave(BpView v, int n_blocks, int shift)
{
    __shared__ __align__(16) BpShared<BZ> sh;
    const int logical = xcd_logical(shift);
    if (logical >= n_blocks) return;
    bp_wave_work<BZ, false>(v, logical, sh);
}


// Fused launch over up to MAX_FUSED instances, one block per logical id in
// the XCD order (as fp_padded_multi; a persistent form cost the mixed
// serving stream 3 % and gained nothing at 512^3).
struct BpMulti {
    BpView v[MAX_FUSED];
    int begin[MAX_FUSED + 1];
    int n;
    int shift;  // chunk size of the XCD order (xcd_logical)
};

// PARTS: instances whose sums are taken in angle parts (BpView::parts > 1),
// run in turn by each block.
template <int BZ, bool PARTS>
__global__ void __launch_bounds__(BR *BR)
__attribute__((amdgpu_waves_per_eu(PARTS ? 1 : bp_waves_per_eu(BZ)))) bp_wave_multi(BpMulti m)
{
    __shared__ __align__(16) BpShared<BZ> sh;
    const int logical = xcd_logical(m.shift);
    if (logical >= m.begin[m.n]) return;
    const int k = owner_of(m.begin, logical);
    bp_wave_work<BZ, PARTS>(m.v[k], logical - m.begin[k], sh);
}

}  // namespace tsx
