// FDK reconstruction kernels (included by tsx.hip); cone-beam plans only.
//
// fdk_filter: one block per (detector row i, angle a). The row is weighted by
//   the cosine of each pixel's ray against the detector normal, staged in LDS
//   between zero margins, and convolved with the sampled Ram-Lak kernel
//   g[0] = 1/4, g[odd m] = -1/(pi^2 m^2) over every lag of the row (the linear
//   convolution of the zero-padded row), scaled by 1 / |U|.
// fdk_bp: one thread per voxel; for every angle the voxel centre is projected
//   from the source onto the virtual detector (FP64), the filtered
//   projections are interpolated bilinearly (zero outside) and accumulated
//   with weight (dbeta / 2) * lambda^2, lambda = D / ((X - s) . n).
#pragma once

#include "tsx_plan.h"
#include "fp_kernels.h"

namespace tsx {

constexpr int FDK_FILTER_THREADS = 128;
constexpr int FDK_FR = 6;   // consecutive outputs per filter thread
constexpr int FDK_FK = 8;   // odd lags per register chunk
constexpr int FDK_MAX_COLS = 4096;
constexpr int FDK_FW = FDK_FR + 2 * FDK_FK;  // window of one side per chunk, floats
constexpr int FDK_PASS = FDK_FILTER_THREADS * FDK_FR;  // outputs per pass over a row

// LDS layout of fdk_filter for a row of `cols` pixels: the weighted row at
// offset fdk_zl(cols) between zeros (wide enough for every lag of every
// output of the last pass), then the lag coefficients.
__host__ __device__ inline int fdk_nodd_pad(int cols)
{
    return (cols / 2 + FDK_FK - 1) / FDK_FK * FDK_FK;
}
__host__ __device__ inline int fdk_zl(int cols) { return 2 * fdk_nodd_pad(cols) + 2; }
__host__ __device__ inline int fdk_lds_floats(int cols)
{
    const int passes = (cols + FDK_PASS - 1) / FDK_PASS;
    return fdk_zl(cols) + passes * FDK_PASS + 2 * fdk_nodd_pad(cols) + FDK_FW + fdk_nodd_pad(cols);
}

constexpr int FDK_BX = 64, FDK_BY = 4;

// With `pair`, q is written row-paired: element (i, a, j) holds the filtered
// values of rows i and i + 1 (0 past the last row), 8 bytes.
//
// The convolution is register-tiled: thread t computes the outputs
// j0 .. j0 + FDK_FR - 1 (j0 = FDK_FR t within a pass) over the odd lags
// m = 2k + 1 in chunks of FDK_FK, from two windows of the row that it reads
// once per chunk (the samples j - m and j + m of all its outputs and lags in
// the chunk), as 8-byte LDS reads; thread starts FDK_FR = 6 floats apart keep
// them free of bank conflicts. One LDS read per output and lag instead of
// three.
//
// The sum is taken in difference form: the kernel has no DC response
// (g[0] + 2 sum of g[odd m] = 0 over all lags), so
//   q[j] = (0.25 - 2 sum_k coef_k) c[j] - sum_k coef_k (c[j - m] + c[j + m] - 2 c[j]),
// with the first factor (`tail`, over the lags summed here) computed in FP64
// by the caller. For a row that is smooth on the pixel scale the plain sum
// nearly cancels 0.25 c[j] and loses about ten times more to FP32 rounding.
__global__ void __launch_bounds__(FDK_FILTER_THREADS)
fdk_filter(const FdkAngle *__restrict__ angles, int n_angles, int rows, int cols,
           const float *__restrict__ proj, float *__restrict__ q, int pair, float tail)
{
    extern __shared__ __align__(16) float lds[];
    const int nodd_pad = fdk_nodd_pad(cols), zl = fdk_zl(cols), nbuf = fdk_lds_floats(cols) - nodd_pad;
    float *buf = lds;           // buf[zl + j]: weighted pixel j; zeros elsewhere
    float *coef = lds + nbuf;   // coef[k] = 1 / (pi^2 (2k + 1)^2)
    const int i = blockIdx.x, a = blockIdx.y;
    const FdkAngle &A = angles[a];
    const int64_t row = ((int64_t)i * n_angles + a) * cols;
    for (int k = threadIdx.x; k < nbuf; k += blockDim.x) {
        const int j = k - zl;
        float v = 0.0f;
        if (j >= 0 && j < cols) {
            const double px = A.P0s[0] + j * A.U[0] + i * A.V[0];
            const double py = A.P0s[1] + j * A.U[1] + i * A.V[1];
            const double pz = A.P0s[2] + j * A.U[2] + i * A.V[2];
            v = (float)(A.D / sqrt(px * px + py * py + pz * pz) * (double)proj[row + j]);
        }
        buf[k] = v;
    }
    for (int k = threadIdx.x; k < nodd_pad; k += blockDim.x) {
        const double m = 2.0 * k + 1.0;
        coef[k] = (float)(1.0 / (M_PI * M_PI * m * m));
    }
    __syncthreads();
    const float inv_du = (float)A.inv_du;
    for (int base = 0; base < cols; base += FDK_PASS) {
        const int j0 = base + FDK_FR * (int)threadIdx.x;  // even
        float part[FDK_FR], c2[FDK_FR];
#pragma unroll
        for (int r = 0; r < FDK_FR; ++r) {
            part[r] = 0.0f;
            c2[r] = 2.0f * buf[zl + j0 + r];
        }
        for (int k0 = 0; k0 < nodd_pad; k0 += FDK_FK) {
            // left window from j0 - 2 (k0 + FDK_FK), right window from j0 + 2 k0 (both even)
            const float2 *lw = reinterpret_cast<const float2 *>(buf + zl + j0 - 2 * (k0 + FDK_FK));
            const float2 *rw = reinterpret_cast<const float2 *>(buf + zl + j0 + 2 * k0);
            float L[FDK_FW], R[FDK_FW], ck[FDK_FK];
#pragma unroll
            for (int x = 0; x < FDK_FW / 2; ++x) {
                const float2 l2 = lw[x], r2 = rw[x];
                L[2 * x] = l2.x;
                L[2 * x + 1] = l2.y;
                R[2 * x] = r2.x;
                R[2 * x + 1] = r2.y;
            }
#pragma unroll
            for (int kk = 0; kk < FDK_FK; ++kk) ck[kk] = coef[k0 + kk];
#pragma unroll
            for (int kk = 0; kk < FDK_FK; ++kk) {
#pragma unroll
                for (int r = 0; r < FDK_FR; ++r) {
                    // samples j - m and j + m, j = j0 + r, m = 2 (k0 + kk) + 1
                    const float sum = (L[r - 2 * kk - 1 + 2 * FDK_FK] + R[r + 2 * kk + 1]) - c2[r];
                    part[r] = fmaf(ck[kk], sum, part[r]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < FDK_FR; ++r) {
            const int j = j0 + r;
            if (j >= cols) break;
            const float v = inv_du * fmaf(tail, buf[zl + j], -part[r]);
            if (!pair) {
                q[row + j] = v;
            } else {
                const int64_t prev = ((int64_t)(i - 1) * n_angles + a) * cols;
                q[2 * (row + j)] = v;
                if (i > 0) q[2 * (prev + j) + 1] = v;
                if (i == rows - 1) q[2 * (row + j) + 1] = 0.0f;
            }
        }
    }
}

__global__ void __launch_bounds__(FDK_BX *FDK_BY)
fdk_bp(const FdkAngle *__restrict__ angles, int n_angles, int rows, int cols,
       const float *__restrict__ q, float *__restrict__ vol, int nx, int ny, int nz, double lo_x,
       double lo_y, double lo_z, double h_x, double h_y, double h_z, int additive)
{
    const int x = blockIdx.x * FDK_BX + threadIdx.x;
    const int y = blockIdx.y * FDK_BY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= nx || y >= ny) return;
    const double X = lo_x + (x + 0.5) * h_x, Y = lo_y + (y + 0.5) * h_y, Z = lo_z + (z + 0.5) * h_z;
    const int64_t astride = cols, rstride = (int64_t)n_angles * cols;
    float acc = 0.0f;
    for (int a = 0; a < n_angles; ++a) {
        const FdkAngle &A = angles[a];
        const double wx = X - A.s[0], wy = Y - A.s[1], wz = Z - A.s[2];
        const double den = wx * A.n[0] + wy * A.n[1] + wz * A.n[2];
        if (!(den > 0.0)) continue;  // voxel not in front of the source
        const double lam = A.D * rcp64(den);
        const float fj = (float)(A.cj + lam * (wx * A.Bj[0] + wy * A.Bj[1] + wz * A.Bj[2]));
        const float fi = (float)(A.ci + lam * (wx * A.Bi[0] + wy * A.Bi[1] + wz * A.Bi[2]));
        if (!(fj > -1.0f && fj < (float)cols && fi > -1.0f && fi < (float)rows)) continue;
        const float fj0 = floorf(fj), fi0 = floorf(fi);
        const int j0 = (int)fj0, i0 = (int)fi0;
        const float wj = fj - fj0, wi = fi - fi0;
        // corners outside the detector read a clamped (in-bounds) pixel and count 0
        const float *base = q + (int64_t)a * astride;
        const bool j_lo = j0 >= 0, j_hi = j0 + 1 < cols, i_lo = i0 >= 0, i_hi = i0 + 1 < rows;
        const int jc0 = max(j0, 0), jc1 = min(j0 + 1, cols - 1);
        const int64_t r0 = (int64_t)max(i0, 0) * rstride, r1 = (int64_t)min(i0 + 1, rows - 1) * rstride;
        const float v00 = (i_lo && j_lo) ? base[r0 + jc0] : 0.0f;
        const float v01 = (i_lo && j_hi) ? base[r0 + jc1] : 0.0f;
        const float v10 = (i_hi && j_lo) ? base[r1 + jc0] : 0.0f;
        const float v11 = (i_hi && j_hi) ? base[r1 + jc1] : 0.0f;
        const float lo = fmaf(wj, v01 - v00, v00);
        const float hi = fmaf(wj, v11 - v10, v10);
        acc = fmaf((float)(A.half_dbeta * lam * lam), fmaf(wi, hi - lo, lo), acc);
    }
    float *o = vol + ((int64_t)z * ny + y) * nx + x;
    *o = additive ? *o + acc : acc;
}

// Upright detectors (v parallel to z, detector normal perpendicular to z, every
// angle; at least 2 detector columns; row-paired q): along a voxel column (x, y) the depth,
// the detector column and the weight are constant and the detector row is
// affine in z. One thread per (x, y) and FDK_ZC voxels along z: FP64 setup
// once per angle, FP32 per voxel; the waves of a block go through the angles
// in step (a barrier per angle).
constexpr int FDK_ZC = 16;

__global__ void __launch_bounds__(FDK_BX *FDK_BY)
fdk_bp_upright(const FdkAngle *__restrict__ angles, int n_angles, int rows, int cols,
               const float *__restrict__ q, float *__restrict__ vol, int nx, int ny, int nz,
               double lo_x, double lo_y, double lo_z, double h_x, double h_y, double h_z, int additive)
{
    const int x = blockIdx.x * FDK_BX + threadIdx.x;
    const int y = blockIdx.y * FDK_BY + threadIdx.y;
    const int z0 = blockIdx.z * FDK_ZC;
    const bool active = x < nx && y < ny;  // every thread takes part in the barriers
    const double X = lo_x + (x + 0.5) * h_x, Y = lo_y + (y + 0.5) * h_y, Z0 = lo_z + (z0 + 0.5) * h_z;
    // the row-paired filtered projections are addressed with 32-bit byte offsets (< 4 GiB)
    const uint32_t rbytes = (uint32_t)n_angles * (uint32_t)cols * 8u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)q, (short)0, (int)(rbytes * (uint32_t)rows), 0x00020000);
    float acc[FDK_ZC];
#pragma unroll
    for (int k = 0; k < FDK_ZC; ++k) acc[k] = 0.0f;
    for (int a = 0; a < n_angles; ++a) {
        auto one_angle = [&]() {
            const FdkAngle &A = angles[a];
            const double wx = X - A.s[0], wy = Y - A.s[1], wz = Z0 - A.s[2];
            const double den = wx * A.n[0] + wy * A.n[1];
            if (!(den > 0.0)) return;
            const double lam = A.D * rcp64(den);
            const float fj = (float)(A.cj + lam * (wx * A.Bj[0] + wy * A.Bj[1]));
            if (!(fj > -1.0f && fj < (float)cols)) return;
            const float fi0 = (float)(A.ci + lam * (wx * A.Bi[0] + wy * A.Bi[1] + wz * A.Bi[2]));
            const float dfi = (float)(lam * h_z * A.Bi[2]);
            const float wgt = (float)(A.half_dbeta * lam * lam);
            const float fj0 = floorf(fj);
            const int j0 = (int)fj0;
            const float wj = fj - fj0;
            // columns (jb, jb + 1), jb = j0 clamped into [0, cols - 2], read as one
            // 8-byte pair; the weights of columns outside the detector are 0
            const int jb = min(max(j0, 0), max(cols - 2, 0));
            float ca, cb;
            if (j0 == jb) {
                ca = 1.0f - wj;
                cb = j0 + 1 < cols ? wj : 0.0f;
            } else if (j0 < jb) {  // j0 = -1: only column 0 = jb, weight wj
                ca = wj;
                cb = 0.0f;
            } else {               // j0 = cols - 1: only column jb + 1, weight 1 - wj
                ca = 0.0f;
                cb = 1.0f - wj;
            }
            // q is row-paired: the 16 bytes at (row i, column jb) are
            // q[i][jb], q[i+1][jb], q[i][jb+1], q[i+1][jb+1]
            const uint32_t cboff = ((uint32_t)a * (uint32_t)cols + (uint32_t)jb) * 8u;
    #pragma unroll
            for (int k = 0; k < FDK_ZC; ++k) {
                const float fi = fmaf((float)k, dfi, fi0);
                const float fi0f = floorf(fi);
                const int i0 = (int)fi0f;
                const float wi = fi - fi0f;
                // rows i0, i0 + 1 from the element of row max(i0, 0) (inside the
                // buffer); for i0 = -1 that element's first row is row i0 + 1
                const uint32_t o = (uint32_t)min(max(i0, 0), rows - 1) * rbytes + cboff;
                const float x0 = bload(rs, o), y0 = bload(rs, o + 4u);
                const float x1 = bload(rs, o + 8u), y1 = bload(rs, o + 12u);
                const float first = fmaf(ca, x0, cb * x1), second = fmaf(ca, y0, cb * y1);
                const bool in = i0 >= 0 && i0 < rows;
                const float top = in ? first : 0.0f;
                const float bot = in ? second : (i0 == -1 ? first : 0.0f);
                acc[k] = fmaf(wgt, fmaf(wi, bot - top, top), acc[k]);
            }
        };
        if (active) one_angle();
        // The block's waves stay at one angle together: their voxels read
        // neighbouring pixels of the same filtered projection, which the
        // waves then share in L1 (17 % faster at 512^3 than free-running).
        __syncthreads();

    }
#pragma unroll
    for (int k = 0; k < FDK_ZC; ++k) {
        const int z = z0 + k;
        if (active && z < nz) {
            float *o = vol + ((int64_t)z * ny + y) * nx + x;
            *o = additive ? *o + acc[k] : acc[k];
        }
    }
}

}  // namespace tsx
