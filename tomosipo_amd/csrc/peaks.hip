// Peak-rate microbenchmarks of the GPU the projectors run on, for the
// roofline figures that bench.py reports (a standalone program, not part of
// libtsx.so):
//   lds_read:   aggregate LDS read bandwidth. 256-thread blocks, every lane
//               reads 16-byte chunks of a 16 KB LDS array (consecutive lanes,
//               consecutive chunks, so no bank conflicts); integer adds
//               consume the data.
//   fp32_fma:   vector FP32 issue rate: 8 independent FMA chains per lane
//               (built without packed FP32, as the projectors are).
//   int_xad:    vector integer issue rate: 8 independent xor-add chains.
// Prints one JSON object per measurement (best of several launches).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int CHUNKS = 1024;  // 16 KB of uint4

__global__ void __launch_bounds__(256) lds_read(uint32_t *out, int iters)
{
    __shared__ uint4 buf[CHUNKS];
    const int t = threadIdx.x;
    for (int i = t; i < CHUNKS; i += 256) buf[i] = make_uint4(i, 3 * i, 5 * i, 7 * i);
    __syncthreads();
    uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int base = t;
    for (int k = 0; k < iters; ++k) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint4 v = buf[(base + u * 256) & (CHUNKS - 1)];
            if (u & 1) {
                a0 = a0 + v.x + v.y;
                a1 = a1 + v.z + v.w;
            } else {
                a2 = a2 + v.x + v.y;
                a3 = a3 + v.z + v.w;
            }
        }
        base = (base + 64) & (CHUNKS - 1);
    }
    out[blockIdx.x * 256 + t] = a0 ^ a1 ^ a2 ^ a3;
}

__global__ void __launch_bounds__(256) fp32_fma(float *out, int iters, float b, float c)
{
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = threadIdx.x * 1e-3f + k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = fmaf(a[k], b, c);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) int_xad(uint32_t *out, int iters, uint32_t b)
{
    uint32_t a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = threadIdx.x + k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = (a[k] ^ b) + (uint32_t)k;  // one v_xad_u32
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s ^= a[k];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <typename F>
float best_ms(F &&launch, hipEvent_t e0, hipEvent_t e1)
{
    launch();  // warm-up
    (void)hipDeviceSynchronize();
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        (void)hipEventRecord(e0, 0);
        launch();
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    return best;
}

}  // namespace

int main()
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
    const int cus = prop.multiProcessorCount;
    int clk_khz = 0;
    (void)hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, 0);
    const double clk = clk_khz * 1e3;
    const int blocks = cus * 32;
    void *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * 4) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    std::printf("{\"device\": \"%s\", \"cus\": %d, \"clock_mhz\": %.0f}\n", prop.gcnArchName, cus, clk / 1e6);

    const int lds_iters = 4096, lds_blocks = cus * 16;
    const float t_lds = best_ms([&] {
        hipLaunchKernelGGL(lds_read, dim3(lds_blocks), dim3(256), 0, 0, (uint32_t *)out, lds_iters);
    }, e0, e1);
    const double lds_bytes = (double)lds_blocks * 256 * lds_iters * 4 * 16;
    std::printf("{\"kernel\": \"lds_read\", \"ms\": %.4f, \"lds_read_TBps\": %.3f, \"bytes_per_cu_clk\": %.1f}\n",
                t_lds, lds_bytes / (t_lds * 1e-3) / 1e12, lds_bytes / (t_lds * 1e-3) / cus / clk);

    const int iters = 2048;
    const double winstr = (double)blocks * 4 * iters * 64;  // wave-wide VALU instructions
    const float t_fma = best_ms([&] {
        hipLaunchKernelGGL(fp32_fma, dim3(blocks), dim3(256), 0, 0, (float *)out, iters, 0.999f, 1e-3f);
    }, e0, e1);
    std::printf("{\"kernel\": \"fp32_fma\", \"ms\": %.4f, \"wave_valu_per_cu_clk\": %.3f, \"fp32_TFLOPs\": %.2f}\n",
                t_fma, winstr / (cus * clk * t_fma * 1e-3), winstr * 64 * 2 / (t_fma * 1e-3) / 1e12);
    const float t_int = best_ms([&] {
        hipLaunchKernelGGL(int_xad, dim3(blocks), dim3(256), 0, 0, (uint32_t *)out, iters, 12345u);
    }, e0, e1);
    std::printf("{\"kernel\": \"int_xad\", \"ms\": %.4f, \"wave_valu_per_cu_clk\": %.3f}\n", t_int,
                winstr / (cus * clk * t_int * 1e-3));
    (void)hipFree(out);
    return 0;
}
