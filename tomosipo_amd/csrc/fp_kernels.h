// Forward-projection kernels (included by tsx.hip).
//
// fp_general: one thread per (sub-)ray, bounds-checked bilinear reads from the
//   caller's [z][y][x] volume; any geometry. Kept as the reference
//   implementation and for volumes whose padded copies would not fit 32-bit
//   byte offsets.
// fp_padded: one thread per (sub-)ray, reading from a zero-padded copy of the
//   volume laid out per stepping axis so that the in-slice axis c2 is
//   contiguous (see the layout kernels below). No per-sample bounds checks; raw buffer
//   loads with 32-bit offsets (out-of-range reads return 0 in hardware).
#pragma once

#include "tsx_plan.h"

namespace tsx {

constexpr int FP_BX = 64, FP_BY = 8;  // FP_BX = 64: one wave per detector row of a block
constexpr int PAD = 2;  // zero border on each side of c1 and c2 in padded layouts

template <typename T>
__device__ __forceinline__ void emit(T *out, int64_t o, int64_t b, T acc, const Epilogue &e)
{
    if (e.mode == EP_SET) {
        out[o] = acc;
    } else if (e.mode == EP_ADD) {
        out[o] += acc;
    } else if (e.mode == EP_RESIDUAL) {
        const int64_t q = b * e.stride + o;
        out[o] = (T)(e.w[q] * (e.y[q] - (float)acc));
    } else if (e.mode == EP_SCALED_ADD) {
        out[o] += (T)(e.w[b * e.stride + o] * (float)acc);
    } else {
        out[o] -= (T)(e.w[b * e.stride + o] * (float)acc);
    }
}

struct RaySetup {
    float a1, a2;    // slope of Q_c per slice
    float R1, R2;    // Q_c at slice nref (padded coordinates in fp_padded)
    float len;       // world length per slice step
    int nref, nlo, nhi;
};

// FP64 setup of the ray through detector point (fj, fi). Returns false when the
// ray has no sample with non-zero weight. Q_c(n) = R_c + (n - nref) a_c.
__device__ __forceinline__ bool ray_setup(const AngleParams &A, double fj, double fi, double off,
                                          RaySetup &r)
{
    const double Pm = A.P0[0] + fj * A.U[0] + fi * A.V[0];
    const double P1 = A.P0[1] + fj * A.U[1] + fi * A.V[1];
    const double P2 = A.P0[2] + fj * A.U[2] + fi * A.V[2];
    const double Dm = A.cone ? Pm - A.S[0] : A.S[0];
    const double D1 = A.cone ? P1 - A.S[1] : A.S[1];
    const double D2 = A.cone ? P2 - A.S[2] : A.S[2];
    if (Dm == 0.0) return false;
    const double inv = 1.0 / Dm;
    const double a1 = D1 * inv, a2 = D2 * inv;
    const int nm = A.n[0], n1 = A.n[1], n2 = A.n[2];
    const int nref = nm / 2;
    const double R1 = P1 + ((double)nref - Pm) * a1;
    const double R2 = P2 + ((double)nref - Pm) * a2;
    // slices n with Q_c(n) in (-1, n_c) for both in-slice axes
    double lo = -1e30, hi = 1e30;
    if (a1 != 0.0) {
        const double t0 = (-1.0 - R1) / a1, t1 = ((double)n1 - R1) / a1;
        lo = fmax(lo, fmin(t0, t1));
        hi = fmin(hi, fmax(t0, t1));
    } else if (!(R1 > -1.0 && R1 < (double)n1)) {
        return false;
    }
    if (a2 != 0.0) {
        const double t0 = (-1.0 - R2) / a2, t1 = ((double)n2 - R2) / a2;
        lo = fmax(lo, fmin(t0, t1));
        hi = fmin(hi, fmax(t0, t1));
    } else if (!(R2 > -1.0 && R2 < (double)n2)) {
        return false;
    }
    lo = fmax(lo + nref, -1.0);
    hi = fmin(hi + nref, (double)nm);
    if (!(lo < hi)) return false;
    r.nlo = max(0, (int)ceil(lo));
    r.nhi = min(nm - 1, (int)floor(hi));
    if (r.nlo > r.nhi) return false;
    r.a1 = (float)a1;
    r.a2 = (float)a2;
    r.R1 = (float)(R1 + off);
    r.R2 = (float)(R2 + off);
    r.len = (float)sqrt(A.h[0] * A.h[0] + A.h[1] * A.h[1] * a1 * a1 + A.h[2] * A.h[2] * a2 * a2);
    r.nref = nref;
    return true;
}

// General forward projector, templated on the data type (float for the
// projector, double for exact gradient checks). Ray setup is FP64 in both.
template <typename T>
__device__ T fp_ray_general(const AngleParams &A, const T *__restrict__ vol, double fj, double fi)
{
    const double Pm = A.P0[0] + fj * A.U[0] + fi * A.V[0];
    const double P1 = A.P0[1] + fj * A.U[1] + fi * A.V[1];
    const double P2 = A.P0[2] + fj * A.U[2] + fi * A.V[2];
    const double Dm = A.cone ? Pm - A.S[0] : A.S[0];
    if (Dm == 0.0) return T(0);
    RaySetup r;
    if (!ray_setup(A, fj, fi, 0.0, r)) return T(0);
    const double inv = 1.0 / Dm;
    const T a1 = (T)(((A.cone ? P1 - A.S[1] : A.S[1])) * inv);
    const T a2 = (T)(((A.cone ? P2 - A.S[2] : A.S[2])) * inv);
    const T R1 = (T)(P1 + ((double)r.nref - Pm) * (double)a1);
    const T R2 = (T)(P2 + ((double)r.nref - Pm) * (double)a2);
    const T len = (T)sqrt(A.h[0] * A.h[0] + A.h[1] * A.h[1] * (double)a1 * (double)a1 +
                          A.h[2] * A.h[2] * (double)a2 * (double)a2);
    const int n1 = A.n[1], n2 = A.n[2];
    const int64_t sm = A.st[0], s1 = A.st[1], s2 = A.st[2];
    T total = T(0), part = T(0);
    int cnt = 0;
    for (int n = r.nlo; n <= r.nhi; ++n) {
        const T dn = (T)(n - r.nref);
        const T q1 = dn * a1 + R1;
        const T q2 = dn * a2 + R2;
        const T f1 = floor(q1), f2 = floor(q2);
        const int k1 = (int)f1, k2 = (int)f2;
        const T w1 = q1 - f1, w2 = q2 - f2;
        const bool a0 = (unsigned)k1 < (unsigned)n1, a1ok = (unsigned)(k1 + 1) < (unsigned)n1;
        const bool b0 = (unsigned)k2 < (unsigned)n2, b1ok = (unsigned)(k2 + 1) < (unsigned)n2;
        const int64_t base = (int64_t)n * sm + (int64_t)k1 * s1 + (int64_t)k2 * s2;
        const T v00 = (a0 && b0) ? vol[base] : T(0);
        const T v10 = (a1ok && b0) ? vol[base + s1] : T(0);
        const T v01 = (a0 && b1ok) ? vol[base + s2] : T(0);
        const T v11 = (a1ok && b1ok) ? vol[base + s1 + s2] : T(0);
        const T lo2 = w1 * (v10 - v00) + v00;
        const T hi2 = w1 * (v11 - v01) + v01;
        part += w2 * (hi2 - lo2) + lo2;
        if (++cnt == 16) {
            total += part;
            part = T(0);
            cnt = 0;
        }
    }
    return len * (total + part);
}

template <typename T>
__global__ void __launch_bounds__(FP_BX *FP_BY)
fp_general(const AngleParams *__restrict__ angles, int n_angles, int a0, int rows, int cols, int ss,
           int tiles_u, const T *__restrict__ vol, int64_t vol_stride, T *__restrict__ proj,
           int64_t proj_stride, Epilogue ep)
{
    // block (tile, angle a0 + blockIdx.y, batch entry blockIdx.z); the host
    // splits angle and batch ranges over launches of at most 65535 each
    const int tile = blockIdx.x;
    const int j = (tile % tiles_u) * FP_BX + threadIdx.x;
    const int i = (tile / tiles_u) * FP_BY + threadIdx.y;
    const int a = a0 + (int)blockIdx.y;
    const int b = blockIdx.z;
    if (j >= cols || i >= rows) return;
    const AngleParams &A = angles[a];
    const T *v = vol + (int64_t)b * vol_stride;
    T acc = T(0);
    const double inv = 1.0 / ss;
    for (int p = 0; p < ss; ++p)
        for (int q = 0; q < ss; ++q)
            acc += fp_ray_general<T>(A, v, j + (q + 0.5) * inv - 0.5, i + (p + 0.5) * inv - 0.5);
    acc *= (T)(inv * inv);
    emit(proj + (int64_t)b * proj_stride, ((int64_t)i * n_angles + a) * cols + j, b, acc, ep);
}

// ---------------------------------------------------------------------------
// padded per-axis layouts
// ---------------------------------------------------------------------------
// Layout for stepping axis x: dims [nz + 2P][nx][ny + 2P], c1 = z, c2 = y.
// Layout for stepping axis y: dims [nz + 2P][ny][nx + 2P], c1 = z, c2 = x.
// Layout for stepping axis z: dims [nz][ny + 2P][nx + 2P], c1 = y, c2 = x.
// Element (m, c1, c2) of an angle sits at m*pst[0] + (c1+P)*pst[1] + (c2+P).

// With `pair`, element (c1, c2) of a layout holds the two floats of rows c1
// and c1 + 1 at c2 (8 bytes; the row past the last is 0), so that the four
// corners of a bilinear sample are 16 contiguous bytes.
//
// y- and z-stepping layouts: thread `xp` of the row (yp, zp) of the layout.
__device__ __forceinline__ void pad_copy_yz_at(const float *__restrict__ src, float *__restrict__ dst, int nx,
                                               int ny, int nz, int axis, int pair, int xp, int yp, int zp)
{
    // axis 1: dst [nz+2P][ny][nx+2P] (c1 = z);  axis 2: dst [nz][ny+2P][nx+2P] (c1 = y)
    const int X = nx + 2 * PAD;
    const int Y = axis == 1 ? ny : ny + 2 * PAD;
    if (xp >= X) return;
    const int x = xp - PAD;
    const int y = axis == 1 ? yp : yp - PAD;
    const int z = axis == 1 ? zp - PAD : zp;
    auto at = [&](int zz, int yy) {
        return (x >= 0 && x < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
                   ? src[((int64_t)zz * ny + yy) * nx + x] : 0.0f;
    };
    const int64_t o = ((int64_t)zp * Y + yp) * X + xp;
    if (pair)
        reinterpret_cast<float2 *>(dst)[o] =
            make_float2(at(z, y), axis == 1 ? at(z + 1, y) : at(z, y + 1));
    else
        dst[o] = at(z, y);
}

// x-stepping layout dst [nz+2P][nx][ny+2P] (c1 = z) from src [nz][ny][nx]:
// the 32 x 32 tile (x0, yp0) of layout plane zp through LDS (two z planes for
// the paired layout), by a block of 32 x 8 threads (tx, ty).
__device__ __forceinline__ void pad_transpose_x_at(const float *__restrict__ src, float *__restrict__ dst,
                                                   int nx, int ny, int nz, int pair, int x0, int yp0, int zp,
                                                   int tx, int ty, float (&tile)[2][32][33])
{
    const int Yp = ny + 2 * PAD;
    for (int d = 0; d <= pair; ++d) {
        const int z = zp - PAD + d;
        for (int r = ty; r < 32; r += 8) {
            const int yp = yp0 + r, x = x0 + tx, y = yp - PAD;
            float v = 0.0f;
            if (z >= 0 && z < nz && y >= 0 && y < ny && x < nx) v = src[((int64_t)z * ny + y) * nx + x];
            tile[d][r][tx] = v;
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int x = x0 + r, yp = yp0 + tx;
        if (x < nx && yp < Yp) {
            const int64_t o = ((int64_t)zp * nx + x) * Yp + yp;
            if (pair)
                reinterpret_cast<float2 *>(dst)[o] = make_float2(tile[0][tx][r], tile[1][tx][r]);
            else
                dst[o] = tile[0][tx][r];
        }
    }
}

// Index k of the job that owns id i (begin[k] <= i < begin[k + 1]): a count
// over every entry, so that the loads of begin[] (kernel arguments) issue
// together rather than one after another as in a search that stops early
// (with 32 layout jobs that search made the launch 1.6x slower). Entries
// past the last job hold INT_MAX.
template <int N>
__device__ __forceinline__ int owner_of(const int (&begin)[N], int i)
{
    int k = 0;
#pragma unroll
    for (int e = 1; e < N - 1; ++e) k += i >= begin[e] ? 1 : 0;
    return k;
}

// The padded layouts of several volumes in one launch: job k (a volume and
// one stepping axis) is row y = k of the grid, whose blocks x < blocks[k]
// build it (32 x 8 threads; the rest of the row returns at once). x layouts
// take (ceil(nx / 32) x ceil((ny + 2P) / 32)) tiles per plane, y and z layouts
// ceil((nx + 2P) / 256) blocks of 256 per row. (With the job found from a
// table of first blocks instead, every block's table loads doubled the time
// of these short blocks.)
constexpr int MAX_LAYOUT_JOBS = 48;

struct LayoutJob {
    const float *src;
    float *dst;
    int nx, ny, nz;
    int axis;  // stepping axis of the layout: 0 (transposed), 1, 2
    int pair;
    int blocks;
};

struct LayoutJobs {
    LayoutJob job[MAX_LAYOUT_JOBS];
    int n;
};

// Elements of the y or z layout of job `j`.
__host__ __device__ inline int64_t layout_job_elems(const LayoutJob &j)
{
    const int64_t X = j.nx + 2 * PAD;
    return j.axis == 1 ? X * j.ny * (j.nz + 2 * PAD) : X * (j.ny + 2 * PAD) * j.nz;
}

// Blocks of job `j` in build_layouts_multi: x layouts a block per 32 x 32
// tile of a plane, y and z layouts a block per 256 consecutive elements of
// the layout (rows of nx + 4 elements rarely fill blocks of their own: at
// 256^3 every row took two blocks, the second with 4 threads busy).
__host__ __device__ inline int64_t layout_job_blocks(const LayoutJob &j)
{
    if (j.axis == 0)
        return (int64_t)((j.nx + 31) / 32) * ((j.ny + 2 * PAD + 31) / 32) * (j.nz + 2 * PAD);
    return (layout_job_elems(j) + 255) / 256;
}

// x layouts (TRANSPOSE, through an LDS tile) or y and z layouts (plain
// copies, no LDS: with the tile allocated for every block the copies ran
// 2.3x slower) of a job list holding only that kind.
template <bool TRANSPOSE>
__global__ void __launch_bounds__(256) build_layouts_multi(LayoutJobs L)
{
    const LayoutJob &J = L.job[blockIdx.y];
    const int r = blockIdx.x;
    if (r >= J.blocks) return;
    if (TRANSPOSE) {
        __shared__ float tile[2][32][33];
        const int tx_tiles = (J.nx + 31) / 32, ty_tiles = (J.ny + 2 * PAD + 31) / 32;
        const int bx = r % tx_tiles, by = (r / tx_tiles) % ty_tiles, zp = r / (tx_tiles * ty_tiles);
        pad_transpose_x_at(J.src, J.dst, J.nx, J.ny, J.nz, J.pair, bx * 32, by * 32, zp, threadIdx.x,
                           threadIdx.y, tile);
    } else {
        const uint32_t X = J.nx + 2 * PAD, Y = J.axis == 1 ? J.ny : J.ny + 2 * PAD;
        const uint32_t e = (uint32_t)r * 256u + threadIdx.y * 32u + threadIdx.x;
        if ((int64_t)e >= layout_job_elems(J)) return;
        const uint32_t row = e / X;
        pad_copy_yz_at(J.src, J.dst, J.nx, J.ny, J.nz, J.axis, J.pair, (int)(e - row * X), (int)(row % Y),
                       (int)(row / Y));
    }
}

// A block-uniform pointer moved to scalar registers. The compiler cannot
// always prove that a pointer derived from the block index is uniform; a
// buffer descriptor built from a pointer it keeps in vector registers turns
// every buffer load into a loop over distinct descriptors.
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T *)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}

struct Layouts {
    const float *ptr[3];   // per stepping axis (x, y, z); null if unused
    uint32_t bytes[3];     // size of each layout in bytes (< 2^32)
    int64_t stride;        // elements between batch entries of the same layout
};

// 1/d to FP64 precision: an FP32 reciprocal as the seed (relative error
// below 3e-7) and one Newton step, which squares the error. Cheaper than an
// FP64 division; d is a ray extent of the order of the volume, far from the
// FP32 range limits.
__device__ __forceinline__ double rcp64(double d)
{
    const double r = (double)__builtin_amdgcn_rcpf((float)d);
    return fma(r, fma(-d, r, 1.0), r);
}

__device__ __forceinline__ int floor_to_int(float q)
{
    int k;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(k) : "v"(q));
    return k;
}

// Corners of one bilinear sample and its two in-slice weights.
struct Corners {
    float v00, v01, v10, v11, w1, w2;
};

__device__ __forceinline__ void request(__amdgpu_buffer_rsrc_t rs, const RaySetup &r, float dn,
                                        uint32_t rowoff, uint32_t rowb, Corners &c)
{
    const float q1 = fmaf(dn, r.a1, r.R1), q2 = fmaf(dn, r.a2, r.R2);
    const uint32_t off = rowoff + __umul24((uint32_t)floor_to_int(q1), rowb) +
                         ((uint32_t)floor_to_int(q2) << 2);
    c.v00 = bload(rs, off);
    c.v01 = bload(rs, off + 4u);
    c.v10 = bload(rs, off + rowb);
    c.v11 = bload(rs, off + rowb + 4u);
    c.w1 = __builtin_amdgcn_fractf(q1);
    c.w2 = __builtin_amdgcn_fractf(q2);
}

__device__ __forceinline__ float interp(const Corners &c)
{
    const float lo = fmaf(c.w2, c.v01 - c.v00, c.v00);
    const float hi = fmaf(c.w2, c.v11 - c.v10, c.v10);
    return fmaf(c.w1, hi - lo, lo);
}

// Marches one ray through a padded layout (Q and k in padded coordinates)
// over the slices where it has support, with four samples in flight per lane:
// a ring of four corner sets, unrolled so that no register copies are needed.
// The slice range keeps Q within [-1, n_c] (up to rounding), inside the zero
// border. Requests run up to three slices past the ray's range (discarded;
// in-layout or zero by the buffer range check).
__device__ __forceinline__ float march_padded_pipe4(__amdgpu_buffer_rsrc_t rs, const RaySetup &r,
                                                    uint32_t st_m, uint32_t st_1)
{
    const uint32_t rowb = st_1 * 4u, stepb = st_m * 4u;
    float dn = (float)(r.nlo - r.nref);
    uint32_t rowoff = (uint32_t)r.nlo * stepb;
    Corners A0, A1, A2, A3;
    request(rs, r, dn, rowoff, rowb, A0);
    request(rs, r, dn + 1.0f, rowoff + stepb, rowb, A1);
    request(rs, r, dn + 2.0f, rowoff + 2u * stepb, rowb, A2);
    request(rs, r, dn + 3.0f, rowoff + 3u * stepb, rowb, A3);
    float s0 = 0.0f, s1 = 0.0f;
    const int cnt = r.nhi - r.nlo + 1;
    for (int k = 0; k < cnt; k += 4) {
        s0 += interp(A0);
        request(rs, r, dn + 4.0f, rowoff + 4u * stepb, rowb, A0);
        const float v1 = interp(A1);
        s1 += (k + 1 < cnt) ? v1 : 0.0f;
        request(rs, r, dn + 5.0f, rowoff + 5u * stepb, rowb, A1);
        const float v2 = interp(A2);
        s0 += (k + 2 < cnt) ? v2 : 0.0f;
        request(rs, r, dn + 6.0f, rowoff + 6u * stepb, rowb, A2);
        const float v3 = interp(A3);
        s1 += (k + 3 < cnt) ? v3 : 0.0f;
        request(rs, r, dn + 7.0f, rowoff + 7u * stepb, rowb, A3);
        dn += 4.0f;
        rowoff += 4u * stepb;
    }
    return r.len * (s0 + s1);
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-synchronous march: every lane of the block samples the same slice n
// at about the same time, over the union of the lanes' slice ranges, so the
// 64 lanes of one detector row read neighbouring voxels of one slice
// (coalesced) and the block's rows read neighbouring volume rows.
// Sample positions are clamped into [lo, hi] = [PAD - 1.5, n_c + PAD + 0.5]:
// positions outside a ray's support land in the zero border and add exactly
// 0, positions inside are unchanged. Four samples in flight per lane.
struct SyncRay {
    float a1, a2, R1, R2, lo1, hi1, lo2, hi2;
};

// Byte offset of row k1: a 24-bit multiply, or (WIDE: rows of 2^24 bytes or
// more) a 32-bit one.
template <bool WIDE>
__device__ __forceinline__ uint32_t row_offset(uint32_t k1, uint32_t rowb)
{
    return WIDE ? k1 * rowb : __umul24(k1, rowb);
}

template <bool WIDE>
__device__ __forceinline__ void request_sync(__amdgpu_buffer_rsrc_t rs, const SyncRay &r, float dn,
                                             uint32_t rowoff, uint32_t rowb, Corners &c)
{
    const float q1 = __builtin_amdgcn_fmed3f(fmaf(dn, r.a1, r.R1), r.lo1, r.hi1);
    const float q2 = __builtin_amdgcn_fmed3f(fmaf(dn, r.a2, r.R2), r.lo2, r.hi2);
    const uint32_t off = rowoff + row_offset<WIDE>((uint32_t)floor_to_int(q1), rowb) +
                         ((uint32_t)floor_to_int(q2) << 2);
    c.v00 = bload(rs, off);
    c.v01 = bload(rs, off + 4u);
    c.v10 = bload(rs, off + rowb);
    c.v11 = bload(rs, off + rowb + 4u);
    c.w1 = __builtin_amdgcn_fractf(q1);
    c.w2 = __builtin_amdgcn_fractf(q2);
}

// Paired layout: the 16 bytes at element (k1, k2) are v[k1][k2], v[k1+1][k2],
// v[k1][k2+1], v[k1+1][k2+1] (one wave-wide 16-byte load per sample).
template <bool WIDE>
__device__ __forceinline__ void request_rowpair(__amdgpu_buffer_rsrc_t rs, const SyncRay &r, float dn,
                                                uint32_t rowoff, uint32_t rowb, Corners &c)
{
    const float q1 = __builtin_amdgcn_fmed3f(fmaf(dn, r.a1, r.R1), r.lo1, r.hi1);
    const float q2 = __builtin_amdgcn_fmed3f(fmaf(dn, r.a2, r.R2), r.lo2, r.hi2);
    const uint32_t off = rowoff + row_offset<WIDE>((uint32_t)floor_to_int(q1), rowb) +
                         ((uint32_t)floor_to_int(q2) << 3);
    c.v00 = bload(rs, off);
    c.v10 = bload(rs, off + 4u);
    c.v01 = bload(rs, off + 8u);
    c.v11 = bload(rs, off + 12u);
    c.w1 = __builtin_amdgcn_fractf(q1);
    c.w2 = __builtin_amdgcn_fractf(q2);
}

// Union of the slice ranges of the block's waves (one wave per detector row).
// All waves then march the same slices, in step: the rays of neighbouring
// detector rows sample neighbouring volume rows, which they then find in L1
// together (with independent waves the 8-row blocks are 12 % slower at 512^3).
template <int BY>
__device__ __forceinline__ void block_range(int &lo, int &hi)
{
    static_assert(FP_BX == 64, "one wave per block row");
    __shared__ int blo[BY], bhi[BY];
    __syncthreads();  // the previous call's reads are done
    if (threadIdx.x == 0) {
        blo[threadIdx.y] = lo;
        bhi[threadIdx.y] = hi;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < BY; ++e) {
        lo = min(lo, blo[e]);
        hi = max(hi, bhi[e]);
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);
}

template <bool ROWPAIR, int BY, bool WIDE = false>
__device__ __forceinline__ float march_padded_sync(__amdgpu_buffer_rsrc_t rs, const RaySetup &r,
                                                   bool valid, int nref, int n1, int n2,
                                                   uint32_t st_m, uint32_t st_1)
{
    int wlo = wave_min(valid ? r.nlo : 0x7fffffff);
    int whi = wave_max(valid ? r.nhi : -1);
    block_range<BY>(wlo, whi);  // block-uniform from here on
    if (wlo > whi) return 0.0f;
    SyncRay q;
    q.lo1 = (float)PAD - 1.5f;
    q.lo2 = q.lo1;
    q.hi1 = (float)(n1 + PAD) + 0.5f;
    q.hi2 = (float)(n2 + PAD) + 0.5f;
    if (valid) {
        q.a1 = r.a1;
        q.a2 = r.a2;
        q.R1 = r.R1;
        q.R2 = r.R2;
    } else {  // a ray without support samples the zero border
        q.a1 = q.a2 = 0.0f;
        q.R1 = q.R2 = q.lo1;
    }
    const uint32_t es = ROWPAIR ? 8u : 4u;  // bytes per layout element
    const uint32_t rowb = st_1 * es, stepb = st_m * es;
    auto req = [&](float dn_, uint32_t ro_, Corners &c_) {
        if (ROWPAIR)
            request_rowpair<WIDE>(rs, q, dn_, ro_, rowb, c_);
        else
            request_sync<WIDE>(rs, q, dn_, ro_, rowb, c_);
    };
    float dn = (float)(wlo - nref);
    uint32_t rowoff = (uint32_t)wlo * stepb;
    Corners A0, A1, A2, A3;
    req(dn, rowoff, A0);
    req(dn + 1.0f, rowoff + stepb, A1);
    req(dn + 2.0f, rowoff + 2u * stepb, A2);
    req(dn + 3.0f, rowoff + 3u * stepb, A3);
    float s0 = 0.0f, s1 = 0.0f;
    const int cnt = whi - wlo + 1;
    for (int k = 0; k < cnt; k += 4) {
        s0 += interp(A0);
        req(dn + 4.0f, rowoff + 4u * stepb, A0);
        const float v1 = interp(A1);
        s1 += (k + 1 < cnt) ? v1 : 0.0f;
        req(dn + 5.0f, rowoff + 5u * stepb, A1);
        __syncthreads();  // the block's waves stay within two slices of each other
        const float v2 = interp(A2);
        s0 += (k + 2 < cnt) ? v2 : 0.0f;
        req(dn + 6.0f, rowoff + 6u * stepb, A2);
        const float v3 = interp(A3);
        s1 += (k + 3 < cnt) ? v3 : 0.0f;
        req(dn + 7.0f, rowoff + 7u * stepb, A3);
        dn += 4.0f;
        rowoff += 4u * stepb;
        __syncthreads();
    }
    return valid ? r.len * (s0 + s1) : 0.0f;
}

// ---------------------------------------------------------------------------
// LDS-staged march (variant 5, plain layouts). The block's rays are nearly
// parallel, so at slice n their samples lie in a small window of the slice
// (about 92 x 13 cells for a 64 x 8 tile at 512^3 / 512 x 768). Per slice the
// block loads that window with coalesced 16-byte loads into LDS (double
// buffered: slice n + 1 is loaded while slice n is sampled, one barrier per
// slice) and every ray reads its four corners from LDS, so the gather leaves
// the vector-memory address path. Samples outside a ray's own slice range are
// skipped (the plain march adds an exact 0 for them), so the result equals
// the plain march's (variant 3) bit for bit: same positions, weights,
// interpolation and summation order (even and odd slices of the block's
// range in two partial sums).
//
// Window. Per ray the sample position q_c(n) = R_c + (n - nref) a_c is
// affine in n; the minimum over the block's rays is concave and the maximum
// convex, so the chords between the block range's end slices bound them from
// below and above at every slice of the range. The chords, widened by one
// cell and clamped to the layout, are kept in 16.16 fixed point, so the
// window origin of each slice is a few scalar integer operations. A block
// whose window would exceed FW_ROWS x FW_W cells at either end of its range
// (the window's size is affine in n) falls back to the plain march.
constexpr int FW_ROWS = 20, FW_W = 128;
constexpr int FW_WP = FW_W + 4;             // LDS row pitch (floats): 16-byte rows, rows 4 banks apart
constexpr int FW_BUF = FW_ROWS * FW_WP;     // floats per window buffer
constexpr int FW_SLOTS = 2;                 // staged 16-byte chunks per thread per slice (at most)
static_assert(FW_ROWS * (FW_W / 4) <= FW_SLOTS * FP_BX * FP_BY, "window chunks per thread");

struct FpLdsShared {
    float win[2][FW_BUF];
    float red[FP_BY][8];  // per wave: min/max of the two in-slice positions at both range ends
};

__device__ __forceinline__ float wave_minf(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_maxf(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// 16.16 fixed-point affine bound x(n) = (a + b (n - n0)) / 2^16 of a chord
// from x0 (slice n0) to x1 (slice n0 + span); `up` rounds it upwards.
struct FxLine {
    int a, b;
};

__device__ __forceinline__ FxLine fx_line(float x0, float x1, int span, bool up)
{
    const double s = span > 0 ? ((double)x1 - (double)x0) / span : 0.0;
    FxLine l;
    l.a = (int)(up ? ceil((double)x0 * 65536.0) : floor((double)x0 * 65536.0));
    l.b = (int)(up ? ceil(s * 65536.0) : floor(s * 65536.0));
    l.a = __builtin_amdgcn_readfirstlane(l.a);
    l.b = __builtin_amdgcn_readfirstlane(l.b);
    return l;
}

// floor (up = false) or ceil of the bound at slice offset dn = n - n0
__device__ __forceinline__ int fx_at(const FxLine &l, int dn, bool up)
{
    const int v = l.a + l.b * dn;
    return up ? (v + 65535) >> 16 : v >> 16;
}

__device__ __forceinline__ float march_lds(__amdgpu_buffer_rsrc_t rs, const RaySetup &r, bool valid, int nref,
                                           int n1, int n2, uint32_t st_m, uint32_t st_1, FpLdsShared &sh)
{
    int wlo = wave_min(valid ? r.nlo : 0x7fffffff);
    int whi = wave_max(valid ? r.nhi : -1);
    block_range<FP_BY>(wlo, whi);  // block-uniform from here on
    if (wlo > whi) return 0.0f;
    const int span = whi - wlo;
    const int w = threadIdx.y, lane = threadIdx.x, t = w * FP_BX + lane;
    // chords of the block's extreme sample positions (padded coordinates)
    {
        const float d0 = (float)(wlo - nref), d1 = (float)(whi - nref);
        const float p10 = valid ? fmaf(d0, r.a1, r.R1) : 1e30f, p11 = valid ? fmaf(d1, r.a1, r.R1) : 1e30f;
        const float p20 = valid ? fmaf(d0, r.a2, r.R2) : 1e30f, p21 = valid ? fmaf(d1, r.a2, r.R2) : 1e30f;
        const float m10 = wave_minf(p10), m11 = wave_minf(p11), m20 = wave_minf(p20), m21 = wave_minf(p21);
        const float x10 = wave_maxf(valid ? p10 : -1e30f), x11 = wave_maxf(valid ? p11 : -1e30f);
        const float x20 = wave_maxf(valid ? p20 : -1e30f), x21 = wave_maxf(valid ? p21 : -1e30f);
        if (lane == 0) {
            sh.red[w][0] = m10;
            sh.red[w][1] = m11;
            sh.red[w][2] = x10;
            sh.red[w][3] = x11;
            sh.red[w][4] = m20;
            sh.red[w][5] = m21;
            sh.red[w][6] = x20;
            sh.red[w][7] = x21;
        }
    }
    __syncthreads();
    float e[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = sh.red[0][k];
#pragma unroll
    for (int q = 1; q < FP_BY; ++q) {
#pragma unroll
        for (int k = 0; k < 8; ++k) e[k] = (k & 2) ? fmaxf(e[k], sh.red[q][k]) : fminf(e[k], sh.red[q][k]);
    }
    // the window lines, widened by one cell
    bool fits = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) fits = fits && fabsf(e[k]) < 30000.0f;
    FxLine lo1 = {0, 0}, hi1 = {0, 0}, lo2 = {0, 0}, hi2 = {0, 0};
    int rows_w = 0, chunks_w = 0;
    if (fits) {
        lo1 = fx_line(e[0] - 1.0f, e[1] - 1.0f, span, false);
        hi1 = fx_line(e[2] + 1.0f, e[3] + 1.0f, span, true);
        lo2 = fx_line(e[4] - 1.0f, e[5] - 1.0f, span, false);
        hi2 = fx_line(e[6] + 1.0f, e[7] + 1.0f, span, true);
        // window size: the distance between the unclamped lines is affine in
        // n, so its largest value is at an end of the range (+1 for
        // rounding); the origins clamped to the layout only shrink it
#pragma unroll
        for (int end = 0; end < 2; ++end) {
            const int dn = end ? span : 0;
            const int o1 = fx_at(lo1, dn, false), u1 = fx_at(hi1, dn, true) + 1;
            const int o2 = fx_at(lo2, dn, false) & ~3, u2 = fx_at(hi2, dn, true) + 1;
            rows_w = max(rows_w, u1 - o1 + 2);
            chunks_w = max(chunks_w, (u2 - o2 + 4) / 4 + 1);
        }
        fits = rows_w <= FW_ROWS && chunks_w * 4 <= FW_W;
    }
    fits = __builtin_amdgcn_readfirstlane((int)fits) != 0;
    if (!fits) return march_padded_sync<false, FP_BY>(rs, r, valid, nref, n1, n2, st_m, st_1);
    rows_w = __builtin_amdgcn_readfirstlane(rows_w);
    chunks_w = __builtin_amdgcn_readfirstlane(chunks_w);
    // staging slots of this thread: chunk c of window row rr, at most FW_SLOTS
    uint32_t soff[FW_SLOTS];  // byte offset in the layout relative to the window origin
    int sdst[FW_SLOTS];       // float index in a window buffer; -1: no slot
#pragma unroll
    for (int k = 0; k < FW_SLOTS; ++k) {
        const int slot = t + k * FP_BX * FP_BY;
        const int rr = slot / chunks_w, c = slot - rr * chunks_w;
        const bool use = rr < rows_w;
        soff[k] = use ? ((uint32_t)rr * st_1 + (uint32_t)c * 4u) * 4u : 0u;
        sdst[k] = use ? rr * FW_WP + c * 4 : -1;
    }
    const uint32_t stepb = st_m * 4u, rowb = st_1 * 4u;
    float4 stage[FW_SLOTS];
    auto origin = [&](int n, int &o1, int &o2) {
        o1 = max(fx_at(lo1, n - wlo, false), 0);
        o2 = max(fx_at(lo2, n - wlo, false), 0) & ~3;
    };
    auto load = [&](int n) {
        int o1, o2;
        origin(n, o1, o2);
        const uint32_t base = (uint32_t)n * stepb + (uint32_t)o1 * rowb + (uint32_t)o2 * 4u;
#pragma unroll
        for (int k = 0; k < FW_SLOTS; ++k)
            if (sdst[k] >= 0)
                stage[k] = __builtin_bit_cast(float4,
                                              __builtin_amdgcn_raw_buffer_load_b128(rs, base + soff[k], 0, 0));
    };
    auto store = [&](int buf) {
        float4 *wb = reinterpret_cast<float4 *>(sh.win[buf]);
#pragma unroll
        for (int k = 0; k < FW_SLOTS; ++k)
            if (sdst[k] >= 0) wb[sdst[k] >> 2] = stage[k];
    };
    // one slice: window of slice n in buffer buf, sample into s
    auto sample = [&](int n, int buf, float dn, float &s) {
        if (valid && n >= r.nlo && n <= r.nhi) {
            int o1, o2;
            origin(n, o1, o2);
            const float q1 = fmaf(dn, r.a1, r.R1), q2 = fmaf(dn, r.a2, r.R2);
            const int k1 = floor_to_int(q1), k2 = floor_to_int(q2);
            // byte offset of cell (k1, k2) in buffer buf: ((k1 - o1) FW_WP + k2 - o2 + buf FW_BUF) * 4
            const uint32_t base = (uint32_t)(o1 * FW_WP + o2 - buf * FW_BUF) * 4u;  // uniform
            const uint32_t a = __umul24((uint32_t)k1, (uint32_t)(FW_WP * 4)) + ((uint32_t)k2 << 2) - base;
            const float *p = reinterpret_cast<const float *>(reinterpret_cast<const char *>(&sh.win[0][0]) + a);
            Corners c;
            c.v00 = p[0];
            c.v01 = p[1];
            c.v10 = p[FW_WP];
            c.v11 = p[FW_WP + 1];
            c.w1 = __builtin_amdgcn_fractf(q1);
            c.w2 = __builtin_amdgcn_fractf(q2);
            s += interp(c);
        }
    };
    float s0 = 0.0f, s1 = 0.0f;  // slices wlo + even and wlo + odd
    load(wlo);
    store(0);
    float dn = (float)(wlo - nref);
    for (int n = wlo; n <= whi; n += 2) {
        // even slice n (buffer 0), then odd slice n + 1 (buffer 1)
        if (n < whi) load(n + 1);
        __syncthreads();  // window of slice n complete; readers of buffer 1 done
        sample(n, 0, dn, s0);
        if (n < whi) store(1);
        if (n + 1 > whi) break;
        if (n + 1 < whi) load(n + 2);
        __syncthreads();  // window of slice n + 1 complete; readers of buffer 0 done
        sample(n + 1, 1, dn + 1.0f, s1);
        if (n + 1 < whi) store(0);
        dn += 2.0f;
    }
    return valid ? r.len * (s0 + s1) : 0.0f;
}

// One instance of the padded forward projector as seen by a block.
struct FpView {
    const AngleParams *angles;
    int n_angles, rows, cols, tiles_u, tiles_v, batch;
    int shear_rows;    // sheared tiles: bound on a lane's offset across the waves (<= tiles_v * BY; 0: none)
    int column_waves;  // 1: a wave's lanes take consecutive rows of a column (tiles_u along rows)
    int a0, na;  // angles a0 .. a0 + na - 1 of the n_angles are projected by this launch
    Layouts L;
    float *proj;
    int64_t proj_stride;
    Epilogue ep;
};

// Work item `logical` of one instance: (batch entry, detector tile, angle),
// angle fastest so that blocks resident together share a detector tile and
// therefore a slab of the volume.
template <int SS, int PIPE, int BY, bool WIDE = false>
__device__ __forceinline__ void fp_padded_work(const FpView &v, int logical)
{
    static_assert(!WIDE || PIPE == 3 || PIPE == 4, "rows of 2^24 bytes or more: the wave-synchronous marches");
    const int a = v.a0 + logical % v.na;
    const int rest = logical / v.na;
    const int tiles = v.tiles_u * v.tiles_v;
    const int tile = rest % tiles;
    const int b = rest / tiles;
    // lanes along a detector row (columns j) or, for column waves, along a
    // column (rows i); the block's BY waves take neighbouring rows (columns)
    const int along = (tile % v.tiles_u) * FP_BX + threadIdx.x;
    int across = (tile / v.tiles_u) * BY + threadIdx.y;
    if (b >= v.batch) return;  // uniform over the block
    const AngleParams &A = v.angles[a];
    if (v.shear_rows) {
        // sheared tiles: the wave's rays follow the detector direction that
        // keeps them on one row of the slices. For a fixed lane position
        // `along` the offset s is fixed, and the positions across the waves
        // are taken modulo the tiled extent R = tiles_v * BY (|s| <= E <= R),
        // so the tiles still cover each pixel exactly once; the waves at the
        // ends hold pixels from both ends.
        const int E = v.shear_rows, R = v.tiles_v * BY;
        const int s = (int)floorf(A.shear * (float)((int)threadIdx.x - FP_BX / 2) + 0.5f);
        across += min(max(s, -E), E);
        across += across < 0 ? R : (across >= R ? -R : 0);
    }
    const int j = v.column_waves ? across : along;
    const int i = v.column_waves ? along : across;
    const bool inside = j < v.cols && i < v.rows;
    if (PIPE < 3 && !inside) return;
    const float *lp = uniform_ptr(v.L.ptr[A.m] + (int64_t)b * v.L.stride);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)lp, (short)0, (int)v.L.bytes[A.m], 0x00020000);
    const uint32_t st_m = (uint32_t)A.pst[0], st_1 = (uint32_t)A.pst[1];
    float acc = 0.0f;
#pragma unroll
    for (int p = 0; p < SS; ++p) {
#pragma unroll
        for (int q = 0; q < SS; ++q) {
            RaySetup r;
            const double fj = SS == 1 ? (double)j : j + (q + 0.5) / SS - 0.5;
            const double fi = SS == 1 ? (double)i : i + (p + 0.5) / SS - 0.5;
            if (PIPE == 5) {
                static_assert(PIPE != 5 || BY == FP_BY, "the LDS-staged march takes 8-row tiles");
                __shared__ __align__(16) FpLdsShared lsh;
                const bool ok = inside && ray_setup(A, fj, fi, (double)PAD, r);
                acc += march_lds(rs, r, ok, A.n[0] / 2, A.n[1], A.n[2], st_m, st_1, lsh);
            } else if (PIPE >= 3) {
                // every lane of the block joins the march (block-wide range,
                // barriers inside; uniform calls)
                const bool ok = inside && ray_setup(A, fj, fi, (double)PAD, r);
                acc += march_padded_sync<PIPE == 4, BY, WIDE>(rs, r, ok, A.n[0] / 2, A.n[1], A.n[2], st_m, st_1);
            } else if (ray_setup(A, fj, fi, (double)PAD, r)) {
                acc += march_padded_pipe4(rs, r, st_m, st_1);
            }
        }
    }
    if (!inside) return;
    if (SS > 1) acc *= 1.0f / (float)(SS * SS);
    emit(v.proj + (int64_t)b * v.proj_stride, ((int64_t)i * v.n_angles + a) * v.cols + j, b, acc,
         v.ep);
}

// Projections of a plan with detector supersampling ss from those of the
// same scan on the detector refined ss times (fine: [rows ss][n_angles][cols
// ss] per batch entry, entries fs floats apart): pixel (i, a, j) is the mean
// of fine pixels (i ss + p, a, j ss + q), p, q < ss, written through the
// epilogue. A block per 256 columns of one (row, angle) line, block ids from
// `first` (launches over ranges of them stay within the dispatch limit);
// blockIdx.z: batch entry.
__global__ void __launch_bounds__(256) reduce_refined(const float *__restrict__ fine, int64_t fs,
                                                       float *__restrict__ out, int64_t os, int rows,
                                                       int n_angles, int cols, int ss, int64_t first,
                                                       Epilogue ep)
{
    const int64_t b = blockIdx.z;
    const int xb = (cols + 255) / 256;
    const int64_t id = first + blockIdx.x;
    const int64_t line = id / xb;  // i * n_angles + a
    const int j = (int)(id % xb) * 256 + (int)threadIdx.x;
    if (j >= cols) return;
    const int i = (int)(line / n_angles), a = (int)(line % n_angles);
    const int64_t fcols = (int64_t)cols * ss;
    const float *src = fine + b * fs + ((int64_t)i * ss * n_angles + a) * fcols + (int64_t)j * ss;
    float acc = 0.0f;
    for (int p = 0; p < ss; ++p)
        for (int q = 0; q < ss; ++q) acc += src[(int64_t)p * n_angles * fcols + q];
    emit(out + b * os, line * cols + j, b, acc * (1.0f / (float)(ss * ss)), ep);
}

// XCD-aware order: consecutive hardware block ids go round-robin to the 8
// XCDs. Logical ids are dealt to the XCDs in chunks of 2^shift consecutive
// ids, chunk k to XCD k % 8, and each XCD runs its chunks in order. The
// blocks resident on one XCD at a time then hold a run of consecutive logical
// ids (neighbouring angles of one detector tile, neighbouring bricks of one
// row), and every XCD holds a share of every part of the launch. With one
// contiguous range per XCD instead, the XCDs that hold the detector's edge
// rows (or the volume's edge slabs) run out of work early.
__device__ __forceinline__ int xcd_logical_of(unsigned vb, int shift)
{
    const int x = (int)(vb % 8u), j = (int)(vb / 8u);
    return ((((j >> shift) << 3) + x) << shift) + (j & ((1 << shift) - 1));
}

__device__ __forceinline__ int xcd_logical(int shift) { return xcd_logical_of(blockIdx.x, shift); }

// Block of FP_BX x BY threads: a tile of BY detector rows (one wave each).
// WIDE: layouts whose rows reach 2^24 bytes (slices of 2048 x 2048 and up),
// row offsets formed with 32-bit multiplies.
template <int SS, int PIPE, int BY, bool WIDE = false>
__global__ void __launch_bounds__(FP_BX *BY) fp_padded(FpView v, int n_blocks, int shift)
{
    const int logical = xcd_logical(shift);
    if (logical >= n_blocks) return;
    fp_padded_work<SS, PIPE, BY, WIDE>(v, logical);
}

// Fused launch over up to MAX_FUSED instances (any shapes): instance k owns
// the logical ids [begin[k], begin[k + 1]).
constexpr int MAX_FUSED = 16;

// One block per logical id in the XCD order, as the single launches. (A
// persistent form, a resident grid claiming the next logical id of its XCD
// from a per-XCD counter, measured slower: 49.4 against 43.6 ms per 512^3
// instance at its best.)
struct FpMulti {
    FpView v[MAX_FUSED];
    int begin[MAX_FUSED + 1];
    int n;
    int shift;  // chunk size of the XCD order (xcd_logical)
};

template <int PIPE, int BY>
__global__ void __launch_bounds__(FP_BX *BY) fp_padded_multi(FpMulti m)  // (rows under 2^24 bytes)
{
    const int logical = xcd_logical(m.shift);
    if (logical >= m.begin[m.n]) return;
    const int k = owner_of(m.begin, logical);
    fp_padded_work<1, PIPE, BY>(m.v[k], logical - m.begin[k]);
}

}  // namespace tsx
