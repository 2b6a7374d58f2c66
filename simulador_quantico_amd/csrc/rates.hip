// Issue-rate microbenchmark of the arithmetic the pass kernels could use:
// FP64 and FP32 vector FMAs (v_fma_f64 / unpacked v_fma_f32) against FP64 and
// FP32 matrix-core MFMAs (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), each
// with enough independent chains per wave and waves per SIMD to hide latency.
// Prints the sustained TFLOP/s of each over a few milliseconds of work.
//
//   make -C simulador_quantico_amd/csrc rates && ./build/hgm/rates
#include <hip/hip_runtime.h>

#include <cstdio>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <class T>
__global__ void __launch_bounds__(256) k_valu(T *out, T c0, T c1, int iters) {
    T a[16];
    const T x = (T)threadIdx.x * (T)1e-6;
    for (int i = 0; i < 16; ++i) a[i] = x + (T)i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if constexpr (sizeof(T) == 4) a[i] = __builtin_fmaf(c0, a[(i + 1) & 15], a[i]);
            else a[i] = __builtin_fma(c0, a[(i + 1) & 15], a[i]);
        }
    }
    T s = 0;
    for (int i = 0; i < 16; ++i) s += a[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s * c1;
}

// The pass kernels' pattern: 16 complex amplitudes per thread in registers,
// a real rotation with one exact 1 per row (out0 = x + t0 * y, out1 = y + t1 * x:
// 4 independent FMAs per pair) on each of the 4 register bits in turn.
template <class T>
__global__ void __launch_bounds__(256) k_pairs(T *out, T t0, T t1, int iters) {
    T re[16], im[16];
    for (int i = 0; i < 16; ++i) {
        re[i] = (T)(threadIdx.x + i) * (T)1e-3;
        im[i] = (T)i * (T)1e-3;
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
#pragma unroll
            for (int k0 = 0; k0 < 16; ++k0) {
                if (k0 & (1 << b)) continue;
                const int k1 = k0 | (1 << b);
                const T xr = re[k0], xi = im[k0], yr = re[k1], yi = im[k1];
                re[k0] = xr + t0 * yr;
                im[k0] = xi + t0 * yi;
                re[k1] = yr + t1 * xr;
                im[k1] = yi + t1 * xi;
            }
        }
    }
    T s = 0;
    for (int i = 0; i < 16; ++i) s += re[i] + im[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_mfma64(double *out, double c, int iters) {
    const double a = c + threadIdx.x * 1e-9, b = c * 0.5;
    d4 acc[4] = {};
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc[0][0] + acc[1][1] + acc[2][2] + acc[3][3];
}

__global__ void __launch_bounds__(256) k_mfma32(float *out, float c, int iters) {
    const float a = c + threadIdx.x * 1e-6f, b = c * 0.5f;
    f4 acc[4] = {};
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc[0][0] + acc[1][1] + acc[2][2] + acc[3][3];
}

int main() {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus < 1) {
        std::printf("no device\n");
        return 1;
    }
    const int blocks = cus * 8, threads = 256;
    void *out = nullptr;
    if (hipMalloc(&out, sizeof(double) * blocks * threads) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const double lanes = (double)blocks * threads, waves = lanes / 64;
    struct Case { const char *name; int iters; double flop_per_iter; };
    const Case cases[6] = {{"FP64 vector FMA (v_fma_f64)", 4096, lanes * 16 * 2},
                           {"FP32 vector FMA (v_fma_f32, unpacked)", 8192, lanes * 16 * 2},
                           {"FP64 MFMA 16x16x4 (v_mfma_f64_16x16x4f64)", 1024, waves * 4 * 16 * 16 * 4 * 2},
                           {"FP32 MFMA 16x16x4 (v_mfma_f32_16x16x4f32)", 2048, waves * 4 * 16 * 16 * 4 * 2},
                           {"FP64 pass-kernel rotations (4 FMA per pair)", 256, lanes * 4 * 8 * 4 * 2},
                           {"FP32 pass-kernel rotations (4 FMA per pair)", 512, lanes * 4 * 8 * 4 * 2}};
    for (int c = 0; c < 6; ++c) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            hipEventRecord(e0);
            const int it = cases[c].iters;
            if (c == 0) hipLaunchKernelGGL(k_valu<double>, dim3(blocks), dim3(threads), 0, 0, (double *)out, 0.999, 1.0, it);
            if (c == 1) hipLaunchKernelGGL(k_valu<float>, dim3(blocks), dim3(threads), 0, 0, (float *)out, 0.999f, 1.0f, it);
            if (c == 2) hipLaunchKernelGGL(k_mfma64, dim3(blocks), dim3(threads), 0, 0, (double *)out, 0.999, it);
            if (c == 3) hipLaunchKernelGGL(k_mfma32, dim3(blocks), dim3(threads), 0, 0, (float *)out, 0.999f, it);
            if (c == 4) hipLaunchKernelGGL(k_pairs<double>, dim3(blocks), dim3(threads), 0, 0, (double *)out, 0.37, -0.21, it);
            if (c == 5) hipLaunchKernelGGL(k_pairs<float>, dim3(blocks), dim3(threads), 0, 0, (float *)out, 0.37f, -0.21f, it);
            hipEventRecord(e1);
            if (hipEventSynchronize(e1) != hipSuccess) return 1;
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0 && ms < best) best = ms;
        }
        std::printf("%-44s %8.3f ms  %7.1f TFLOP/s\n", cases[c].name, best,
                    cases[c].flop_per_iter * cases[c].iters / (best * 1e-3) / 1e12);
    }
    hipFree(out);
    return 0;
}
