// Per-angle projector parameters shared by the host plan builder and the HIP
// kernels. All vectors are stored in the angle's own axis order
// (m, c1, c2): m is the stepping (dominant) axis, c1 and c2 the two in-slice
// axes, c2 being the one with the smaller memory stride.
#pragma once

#include <stdint.h>

namespace tsx {

struct AngleParams {
    // Ray family, grid units (voxel centres on integers). Pixel (i, j) maps to
    // the point P0 + j*U + i*V of a virtual detector plane through the volume
    // centre that is parallel to the real detector; rays are unchanged.
    double P0[3], U[3], V[3];
    double S[3];          // cone: source position; parallel: ray direction
    double N[3];          // U x V
    double Bj[3], Bi[3];  // dual basis: (X - P0) . Bj = j, (X - P0) . Bi = i for X on the plane
    double kappa;         // cone: (P0 - S) . N
    double EN;            // parallel: S . N
    double h[3];          // voxel size along (m, c1, c2)
    int32_t m, c1, c2;    // axis ids, 0 = x, 1 = y, 2 = z
    int32_t cone;
    int32_t n[3];         // volume extent along (m, c1, c2)
    int32_t pad;
    int64_t st[3];        // element strides along (m, c1, c2) in the [z][y][x] volume
    int64_t pst[3];       // element strides along (m, c1, c2) in the padded layout of axis m
    // per-angle constants of the general backprojector
    double inv_kappa;     // cone: 1 / kappa
    double inv_EN;        // parallel: 1 / EN
    double inv_S0;        // parallel: 1 / S[0] (the ray's component along m)
    float fU[3], fV[3];   // U and V rounded to float
    // parallel beam: detector position of grid point k = (k_m, k_c1, k_c2),
    // j = par_c[0] + k . par_c[1..3], i = par_c[4] + k . par_c[5..7]; the
    // local map's terms (A1j, A1i, A2j, A2i) and the half-widths (in
    // pixels) of the image of a voxel's +-1 square, par_b[0] < 0 if the map
    // is singular
    double par_c[8];
    float par_a[4], par_b[2];
    // forward tiles: slope of the detector direction along which rays meet
    // one c1 row of the slices, di/dj for waves along detector rows (0: the
    // rows themselves), dj/di for waves along columns; a wave's lane tx takes
    // the offset round(shear (tx - 32)) across the waves
    float shear;
};

// Per-angle FDK quantities of a cone-beam plan, world (x, y, z) coordinates.
// The detector is moved to the parallel plane through the volume centre c
// (virtual detector: pixel (i, j) centre at s + P0s + j*U + i*V).
struct FdkAngle {
    double s[3];          // source
    double n[3];          // unit detector normal, pointing away from the source
    double P0s[3];        // virtual centre of pixel (0, 0) minus the source
    double U[3], V[3];    // virtual column and row steps
    double Bj[3], Bi[3];  // dual basis on the virtual plane
    double D;             // (c - s) . n
    double cj, ci;        // pixel coordinates (j, i) of the source's foot point: -P0s . Bj, -P0s . Bi
    double half_dbeta;    // half the angular step of the source around the z axis through c
    double inv_du;        // 1 / |U|
};

// How a kernel writes its result `acc` to element o of batch entry b.
struct Epilogue {
    int32_t mode;    // 0: out = acc; 1: out += acc; 2: out = w * (y - acc); 3: out += w * acc;
                     // 4: out -= w * acc (later chunks of a chunked mode-2 launch)
    const float *y;  // mode 2: data term, same layout as the output
    const float *w;  // modes 2 and 3: elementwise weight, same layout as the output
    int64_t stride;  // elements between batch entries of y and w
};

enum { EP_SET = 0, EP_ADD = 1, EP_RESIDUAL = 2, EP_SCALED_ADD = 3, EP_SCALED_SUB = 4 };

}  // namespace tsx
