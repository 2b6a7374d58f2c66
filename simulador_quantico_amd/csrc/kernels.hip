// gfx950 kernels of the t_GPU path.
//
// k_pass_lds: one workgroup per tile of 2^r amplitudes. The tile's local index
// L maps to the state index base | (L & cmask) | deposit(L >> c), where the
// low c bits stay contiguous (so a wave's loads cover runs of 2^c amplitudes)
// and deposit() scatters the remaining r-c tile bits onto the pass's chosen
// state-index bits through three 64-entry LDS tables. Thread t owns local
// indices t + k*B (k < 2R) for loading, storing and diagonal gates; for a
// pair-mixing gate it takes pairs t + k*B (k < R) and inserts a zero at the
// target's tile bit. Barriers are placed only where ownership changes.
#include "kernels.hpp"

namespace hgm {

namespace {

template <class T> struct Vec;
template <> struct Vec<double> { using type = double2; };
template <> struct Vec<float> { using type = float2; };

__device__ __forceinline__ uint64_t pdep64(uint64_t v, uint64_t mask) {
    uint64_t r = 0;
    for (uint64_t bb = 1; mask; bb <<= 1) {
        uint64_t lowbit = mask & (~mask + 1);
        if (v & bb) r |= lowbit;
        mask &= mask - 1;
    }
    return r;
}

template <class V, class T>
__device__ __forceinline__ V cmul(T mr, T mi, V a) {
    V o;
    o.x = mr * a.x - mi * a.y;
    o.y = mr * a.y + mi * a.x;
    return o;
}

template <class V, class T>
__device__ __forceinline__ V cmul2(T ar, T ai, V x, T br, T bi, V y) {
    V o;
    o.x = ar * x.x - ai * x.y + br * y.x - bi * y.y;
    o.y = ar * x.y + ai * x.x + br * y.y + bi * y.x;
    return o;
}

template <class T>
__global__ void __launch_bounds__(1024) k_pass_lds(typename Vec<T>::type *__restrict__ state,
                                                   const DevOp *__restrict__ ops, int n_ops, PassArgs a) {
    using V = typename Vec<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *dep = reinterpret_cast<uint64_t *>(smem);      // 3 tables of 64 entries
    V *tile = reinterpret_cast<V *>(smem + 3 * 64 * sizeof(uint64_t));

    const int t = threadIdx.x;
    const int hi_bits = a.r - a.c;
    for (int x = t; x < 192; x += blockDim.x) {
        const int tab = x >> 6, v = x & 63;
        uint64_t acc = 0;
        for (int j = 0; j < 6; ++j) {
            const int lb = tab * 6 + j;
            if (((v >> j) & 1) && lb < hi_bits) acc |= uint64_t(1) << a.local_pos[a.c + lb];
        }
        dep[x] = acc;
    }
    __syncthreads();

    const int B = a.B, R = a.R, c = a.c;
    const uint32_t cmask = (1u << c) - 1u;
    for (int64_t tid = blockIdx.x; tid < a.n_tiles; tid += gridDim.x) {
        const uint64_t base = pdep64((uint64_t)tid, a.nonlocal_mask);
        // load
        for (int k = 0; k < 2 * R; ++k) {
            const uint32_t L = (uint32_t)(t + k * B);
            const uint32_t h = L >> c;
            const uint64_t g = base | (L & cmask) | dep[h & 63] | dep[64 + ((h >> 6) & 63)] | dep[128 + ((h >> 12) & 63)];
            tile[L] = state[g];
        }
        bool last_pair = false;
        for (int i = 0; i < n_ops; ++i) {
            const DevOp &o = ops[i];
            if ((base & o.gmask) != o.gval) continue;   // tile-uniform
            const T m0r = (T)o.m[0], m0i = (T)o.m[1], m1r = (T)o.m[2], m1i = (T)o.m[3];
            const T m2r = (T)o.m[4], m2i = (T)o.m[5], m3r = (T)o.m[6], m3i = (T)o.m[7];
            if (o.kind == K_DENSE || o.kind == K_ANTI) {
                __syncthreads();
                const int tb = o.tbit;
                const uint32_t tmask = 1u << tb;
                for (int k = 0; k < R; ++k) {
                    const uint32_t p = (uint32_t)(t + k * B);
                    const uint32_t p0 = ((p >> tb) << (tb + 1)) | (p & (tmask - 1u));
                    const uint32_t p1 = p0 | tmask;
                    if ((p0 & o.lmask) != o.lval) continue;
                    const V x = tile[p0], y = tile[p1];
                    if (o.kind == K_DENSE) {
                        tile[p0] = cmul2<V, T>(m0r, m0i, x, m1r, m1i, y);
                        tile[p1] = cmul2<V, T>(m2r, m2i, x, m3r, m3i, y);
                    } else {
                        tile[p0] = cmul<V, T>(m1r, m1i, y);
                        tile[p1] = cmul<V, T>(m2r, m2i, x);
                    }
                }
                last_pair = true;
            } else {
                if (last_pair) { __syncthreads(); last_pair = false; }
                if (o.kind == K_DIAG) {
                    const int tb = o.tbit;
                    for (int k = 0; k < 2 * R; ++k) {
                        const uint32_t L = (uint32_t)(t + k * B);
                        if ((L & o.lmask) != o.lval) continue;
                        const bool one = (L >> tb) & 1u;
                        tile[L] = one ? cmul<V, T>(m3r, m3i, tile[L]) : cmul<V, T>(m0r, m0i, tile[L]);
                    }
                } else {   // K_DIAG_GLOBAL
                    const bool one = (base >> o.tbit) & 1u;
                    const T pr = one ? m3r : m0r, pi = one ? m3i : m0i;
                    for (int k = 0; k < 2 * R; ++k) {
                        const uint32_t L = (uint32_t)(t + k * B);
                        if ((L & o.lmask) != o.lval) continue;
                        tile[L] = cmul<V, T>(pr, pi, tile[L]);
                    }
                }
            }
        }
        if (last_pair) __syncthreads();
        // store
        for (int k = 0; k < 2 * R; ++k) {
            const uint32_t L = (uint32_t)(t + k * B);
            const uint32_t h = L >> c;
            const uint64_t g = base | (L & cmask) | dep[h & 63] | dep[64 + ((h >> 6) & 63)] | dep[128 + ((h >> 12) & 63)];
            state[g] = tile[L];
        }
        __syncthreads();   // the next tile reuses the LDS
    }
}

template <class T>
__global__ void k_set_basis(typename Vec<T>::type *state, uint64_t n, uint64_t index) {
    using V = typename Vec<T>::type;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        V v;
        v.x = (i == index) ? T(1) : T(0);
        v.y = T(0);
        state[i] = v;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_prob_halves(const typename Vec<T>::type *state, uint64_t n, int bit,
                                                      double *partial) {
    __shared__ double s0[256], s1[256];
    double a0 = 0, a1 = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const auto v = state[i];
        const double p = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        if (bit >= 0 && ((i >> bit) & 1)) a1 += p; else a0 += p;
    }
    s0[threadIdx.x] = a0;
    s1[threadIdx.x] = a1;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s0[threadIdx.x] += s0[threadIdx.x + w]; s1[threadIdx.x] += s1[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = s0[0]; partial[2 * blockIdx.x + 1] = s1[0]; }
}

template <class T>
__global__ void __launch_bounds__(256) k_argmax(const typename Vec<T>::type *state, uint64_t n, double *pmax,
                                                 uint64_t *pidx) {
    __shared__ double sm[256];
    __shared__ uint64_t si[256];
    double best = -1.0;
    uint64_t bi = ~uint64_t(0);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const auto v = state[i];
        const double p = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        if (p > best) { best = p; bi = i; }
    }
    sm[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = sm[threadIdx.x + w];
            const uint64_t oi = si[threadIdx.x + w];
            if (o > sm[threadIdx.x] || (o == sm[threadIdx.x] && oi < si[threadIdx.x])) { sm[threadIdx.x] = o; si[threadIdx.x] = oi; }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { pmax[blockIdx.x] = sm[0]; pidx[blockIdx.x] = si[0]; }
}

template <class T>
__global__ void k_collapse(typename Vec<T>::type *state, uint64_t half, int bit, int outcome, double norm) {
    using V = typename Vec<T>::type;
    const uint64_t low = (uint64_t(1) << bit) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < half; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t pos = ((i >> bit) << (bit + 1)) | (i & low);
        const V v = state[pos | (uint64_t(outcome) << bit)];
        V w;
        w.x = (T)((double)v.x / norm);
        w.y = (T)((double)v.y / norm);
        state[pos] = w;
        V z;
        z.x = T(0);
        z.y = T(0);
        state[pos | (uint64_t(1) << bit)] = z;
    }
}

int grid_for(uint64_t n, int block) {
    uint64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

template <class T>
hipError_t pass_lds(void *state, const DevOp *ops, int n_ops, const PassArgs &a, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipFuncSetAttribute((const void *)k_pass_lds<T>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    const size_t lds = 3 * 64 * sizeof(uint64_t) + (size_t(1) << a.r) * sizeof(typename Vec<T>::type);
    int64_t grid = a.n_tiles < (int64_t(1) << 22) ? a.n_tiles : (int64_t(1) << 22);
    hipLaunchKernelGGL(k_pass_lds<T>, dim3((unsigned)grid), dim3(a.B), lds, s,
                       reinterpret_cast<typename Vec<T>::type *>(state), ops, n_ops, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_pass_lds(int dtype, void *state, const DevOp *ops, int n_ops, const PassArgs &a, hipStream_t s) {
    return dtype == HGM_C64 ? pass_lds<float>(state, ops, n_ops, a, s) : pass_lds<double>(state, ops, n_ops, a, s);
}

hipError_t launch_set_basis(int dtype, void *state, uint64_t n, uint64_t index, hipStream_t s) {
    const int blk = 256, grid = grid_for(n, blk);
    if (dtype == HGM_C64) hipLaunchKernelGGL(k_set_basis<float>, dim3(grid), dim3(blk), 0, s, (float2 *)state, n, index);
    else hipLaunchKernelGGL(k_set_basis<double>, dim3(grid), dim3(blk), 0, s, (double2 *)state, n, index);
    return hipGetLastError();
}

hipError_t launch_prob_halves(int dtype, const void *state, uint64_t n, int bit, double *partial, int blocks,
                              hipStream_t s) {
    if (dtype == HGM_C64)
        hipLaunchKernelGGL(k_prob_halves<float>, dim3(blocks), dim3(256), 0, s, (const float2 *)state, n, bit, partial);
    else
        hipLaunchKernelGGL(k_prob_halves<double>, dim3(blocks), dim3(256), 0, s, (const double2 *)state, n, bit, partial);
    return hipGetLastError();
}

hipError_t launch_argmax(int dtype, const void *state, uint64_t n, double *pmax, uint64_t *pidx, int blocks,
                         hipStream_t s) {
    if (dtype == HGM_C64)
        hipLaunchKernelGGL(k_argmax<float>, dim3(blocks), dim3(256), 0, s, (const float2 *)state, n, pmax, pidx);
    else
        hipLaunchKernelGGL(k_argmax<double>, dim3(blocks), dim3(256), 0, s, (const double2 *)state, n, pmax, pidx);
    return hipGetLastError();
}

hipError_t launch_collapse(int dtype, void *state, uint64_t n, int bit, int outcome, double norm, hipStream_t s) {
    const uint64_t half = n / 2;
    const int blk = 256, grid = grid_for(half, blk);
    if (dtype == HGM_C64)
        hipLaunchKernelGGL(k_collapse<float>, dim3(grid), dim3(blk), 0, s, (float2 *)state, half, bit, outcome, norm);
    else
        hipLaunchKernelGGL(k_collapse<double>, dim3(grid), dim3(blk), 0, s, (double2 *)state, half, bit, outcome, norm);
    return hipGetLastError();
}

}  // namespace hgm
