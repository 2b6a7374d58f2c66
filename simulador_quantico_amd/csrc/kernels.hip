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

// State accesses of a single-circuit pass, which reads and writes every
// amplitude once per pass: nontemporal (streaming) hint (interpreted random30
// FP64 609 -> 578 ms per step, Hadamard x3 at 30 qubits 53.6 -> 52.7 ms, on
// one box). The batched serving kernel keeps plain accesses: many of its
// states fit the last-level cache between steps.
template <class T>
__device__ __forceinline__ typename Vec<T>::type load_stream(const typename Vec<T>::type *p) {
    typedef T N __attribute__((ext_vector_type(2)));
    const N v = __builtin_nontemporal_load(reinterpret_cast<const N *>(p));
    return {v.x, v.y};
}
template <class T>
__device__ __forceinline__ void store_stream(typename Vec<T>::type *p, const typename Vec<T>::type &a) {
    typedef T N __attribute__((ext_vector_type(2)));
    const N v = {a.x, a.y};
    __builtin_nontemporal_store(v, reinterpret_cast<N *>(p));
}

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

// ---------------------------------------------------------------------------
// k_pass_reg: the register kernel. Each thread keeps A amplitudes of its tile
// in registers for the whole pass. Gates act on register pairs only (the
// planner's second level schedules them in phases); between phases the tile
// makes one round trip through LDS into the next phase's bit assignment.
// HBM traffic per pass: one read and one write of every amplitude.
//
// Every per-gate condition is resolved by selecting the operator, never by
// branching per amplitude: a register-bit condition picks the matrix or the
// identity per register pair (tile-uniform, scalar selects), a thread-bit
// condition picks it per lane once per gate; the updates themselves are
// straight-line FMA code.

template <class T> struct Mat {
    T m0r, m0i, m1r, m1i, m2r, m2i, m3r, m3i;
};

template <class T>
__device__ __forceinline__ Mat<T> load_mat(const RegOp &o) {
    return {(T)o.m[0], (T)o.m[1], (T)o.m[2], (T)o.m[3], (T)o.m[4], (T)o.m[5], (T)o.m[6], (T)o.m[7]};
}

template <class T>
__device__ __forceinline__ Mat<T> or_identity(const Mat<T> &m, bool id) {
    return {id ? T(1) : m.m0r, id ? T(0) : m.m0i, id ? T(0) : m.m1r, id ? T(0) : m.m1i,
            id ? T(0) : m.m2r, id ? T(0) : m.m2i, id ? T(1) : m.m3r, id ? T(0) : m.m3i};
}

template <class T, int A, int BIT, bool COND>
__device__ __forceinline__ void reg_dense(typename Vec<T>::type (&a)[A], const Mat<T> &m, uint32_t rmask,
                                          uint32_t rval) {
    using V = typename Vec<T>::type;
#pragma unroll
    for (int k0 = 0; k0 < A; ++k0) {
        if (k0 & (1 << BIT)) continue;
        const int k1 = k0 | (1 << BIT);
        const Mat<T> u = COND ? or_identity(m, ((uint32_t)k0 & rmask) != rval) : m;
        const V x = a[k0], y = a[k1];
        a[k0] = cmul2<V, T>(u.m0r, u.m0i, x, u.m1r, u.m1i, y);
        a[k1] = cmul2<V, T>(u.m2r, u.m2i, x, u.m3r, u.m3i, y);
    }
}

// Real 2x2 matrix (imaginary parts zero): half the multiply-adds of reg_dense.
template <class T, int A, int BIT, bool COND>
__device__ __forceinline__ void reg_dense_real(typename Vec<T>::type (&a)[A], const Mat<T> &m, uint32_t rmask,
                                               uint32_t rval) {
    using V = typename Vec<T>::type;
#pragma unroll
    for (int k0 = 0; k0 < A; ++k0) {
        if (k0 & (1 << BIT)) continue;
        const int k1 = k0 | (1 << BIT);
        const bool id = COND && (((uint32_t)k0 & rmask) != rval);
        const T u0 = id ? T(1) : m.m0r, u1 = id ? T(0) : m.m1r, u2 = id ? T(0) : m.m2r, u3 = id ? T(1) : m.m3r;
        const V x = a[k0], y = a[k1];
        V p, q;
        p.x = u0 * x.x + u1 * y.x;
        p.y = u0 * x.y + u1 * y.y;
        q.x = u2 * x.x + u3 * y.x;
        q.y = u2 * x.y + u3 * y.y;
        a[k0] = p;
        a[k1] = q;
    }
}

// dispatch on the runtime register bit (tile-uniform) to unrolled code. With
// STRICT, the cases no bit can reach (the planner emits register bits below
// log2(A) only) are marked unreachable: every remaining case then updates the
// amplitudes in place, and the compiler stops copying them to a second
// register set per gate, for a few more VGPRs (used by the batched kernel,
// whose register budget allows it)
#define HGM_NO_CASE                               \
    if constexpr (STRICT) __builtin_unreachable(); \
    break;
#define HGM_REG_BITS(CALL)                                             \
    switch (bit) {                                                     \
        case 0: CALL(0); break;                                        \
        case 1: if constexpr (A > 2) { CALL(1); break; } HGM_NO_CASE   \
        case 2: if constexpr (A > 4) { CALL(2); break; } HGM_NO_CASE   \
        case 3: if constexpr (A > 8) { CALL(3); break; } HGM_NO_CASE   \
        case 4: if constexpr (A > 16) { CALL(4); break; } HGM_NO_CASE  \
        default: HGM_NO_CASE                                           \
    }

// Rows with an exact unit entry (R_DENSE_PIV*): out_i = x[P_i] + t_i * x[1 - P_i],
// 8 multiply-adds per pair, 4 when t0 and t1 are real.
template <class T, int A, int BIT, int P0, int P1, bool REAL>
__device__ __forceinline__ void reg_dense_piv(typename Vec<T>::type (&a)[A], T t0r, T t0i, T t1r, T t1i) {
    using V = typename Vec<T>::type;
#pragma unroll
    for (int k0 = 0; k0 < A; ++k0) {
        if (k0 & (1 << BIT)) continue;
        const int k1 = k0 | (1 << BIT);
        const V x[2] = {a[k0], a[k1]};
        V p, q;
        if constexpr (REAL) {
            p.x = x[P0].x + t0r * x[1 - P0].x;
            p.y = x[P0].y + t0r * x[1 - P0].y;
            q.x = x[P1].x + t1r * x[1 - P1].x;
            q.y = x[P1].y + t1r * x[1 - P1].y;
        } else {
            p.x = x[P0].x + t0r * x[1 - P0].x - t0i * x[1 - P0].y;
            p.y = x[P0].y + t0r * x[1 - P0].y + t0i * x[1 - P0].x;
            q.x = x[P1].x + t1r * x[1 - P1].x - t1i * x[1 - P1].y;
            q.y = x[P1].y + t1r * x[1 - P1].y + t1i * x[1 - P1].x;
        }
        a[k0] = p;
        a[k1] = q;
    }
}

template <class T, int A, int P0, int P1, bool REAL, bool STRICT>
__device__ __forceinline__ void reg_dense_piv_bits(typename Vec<T>::type (&a)[A], int bit, T t0r, T t0i, T t1r,
                                                   T t1i) {
#define HGM_CALL(B) reg_dense_piv<T, A, B, P0, P1, REAL>(a, t0r, t0i, t1r, t1i)
    HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
}

template <class T, int A, bool REAL, bool STRICT>
__device__ __forceinline__ void reg_dense_piv_any(typename Vec<T>::type (&a)[A], const RegOp &o) {
    const T t0r = (T)o.m[0], t0i = (T)o.m[1], t1r = (T)o.m[2], t1i = (T)o.m[3];
    switch (o.tval & 3u) {
        case 0: reg_dense_piv_bits<T, A, 0, 0, REAL, STRICT>(a, o.rbit, t0r, t0i, t1r, t1i); break;
        case 1: reg_dense_piv_bits<T, A, 1, 0, REAL, STRICT>(a, o.rbit, t0r, t0i, t1r, t1i); break;
        case 2: reg_dense_piv_bits<T, A, 0, 1, REAL, STRICT>(a, o.rbit, t0r, t0i, t1r, t1i); break;
        default: reg_dense_piv_bits<T, A, 1, 1, REAL, STRICT>(a, o.rbit, t0r, t0i, t1r, t1i); break;
    }
}

template <class T, int A, int BIT>
__device__ __forceinline__ void reg_anti(typename Vec<T>::type (&a)[A], const Mat<T> &m) {
    using V = typename Vec<T>::type;
#pragma unroll
    for (int k0 = 0; k0 < A; ++k0) {
        if (k0 & (1 << BIT)) continue;
        const int k1 = k0 | (1 << BIT);
        const V x = a[k0], y = a[k1];
        a[k0] = cmul<V, T>(m.m1r, m.m1i, y);
        a[k1] = cmul<V, T>(m.m2r, m.m2i, x);
    }
}

template <class T, int A, int BIT>
__device__ __forceinline__ void reg_swap(typename Vec<T>::type (&a)[A]) {
    using V = typename Vec<T>::type;
#pragma unroll
    for (int k0 = 0; k0 < A; ++k0) {
        if (k0 & (1 << BIT)) continue;
        const int k1 = k0 | (1 << BIT);
        const V x = a[k0];
        a[k0] = a[k1];
        a[k1] = x;
    }
}

template <class T, int A, bool STRICT>
__device__ __forceinline__ void reg_gate(typename Vec<T>::type (&a)[A], const RegOp &o, int t) {
    const int bit = o.rbit;
    if (o.kind == R_DENSE_PIV) { reg_dense_piv_any<T, A, false, STRICT>(a, o); return; }
    if (o.kind == R_DENSE_PIV_REAL) { reg_dense_piv_any<T, A, true, STRICT>(a, o); return; }
    if (o.kind == R_SWAP) {
#define HGM_CALL(B) reg_swap<T, A, B>(a)
        HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
        return;
    }
    Mat<T> m = load_mat<T>(o);
    if (o.kind == R_ANTI) {
#define HGM_CALL(B) reg_anti<T, A, B>(a, m)
        HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
        return;
    }
    if (o.kind == R_DENSE_REAL) {
        if (o.tmask) m = or_identity(m, ((uint32_t)t & o.tmask) != o.tval);   // per lane, once per gate
        if (o.rmask) {
#define HGM_CALL(B) reg_dense_real<T, A, B, true>(a, m, o.rmask, o.rval)
            HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
        } else {
#define HGM_CALL(B) reg_dense_real<T, A, B, false>(a, m, 0u, 0u)
            HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
        }
        return;
    }
    if (o.tmask == 0) {   // tile-uniform matrix (scalar registers)
        if (o.rmask) {
#define HGM_CALL(B) reg_dense<T, A, B, true>(a, m, o.rmask, o.rval)
            HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
        } else {
#define HGM_CALL(B) reg_dense<T, A, B, false>(a, m, 0u, 0u)
            HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
        }
        return;
    }
    m = or_identity(m, ((uint32_t)t & o.tmask) != o.tval);   // per lane, once per gate
#define HGM_CALL(B) reg_dense<T, A, B, true>(a, m, o.rmask, o.rval)
    HGM_REG_BITS(HGM_CALL)
#undef HGM_CALL
}

template <class T, int A>
__device__ __forceinline__ void reg_phase_apply(typename Vec<T>::type (&a)[A], T fr, T fi, bool real,
                                                uint32_t rmask, uint32_t rval) {
    if (real) {
#pragma unroll
        for (int k = 0; k < A; ++k) {
            const T fk = (((uint32_t)k & rmask) == rval) ? fr : T(1);
            a[k].x *= fk;
            a[k].y *= fk;
        }
    } else {
#pragma unroll
        for (int k = 0; k < A; ++k) {
            const bool hit = ((uint32_t)k & rmask) == rval;
            a[k] = cmul<typename Vec<T>::type, T>(hit ? fr : T(1), hit ? fi : T(0), a[k]);
        }
    }
}

template <class T, int A>
__device__ __forceinline__ void reg_phase(typename Vec<T>::type (&a)[A], const RegOp &o, int t) {
    const bool real = o.kind == R_PHASE_REAL;
    if (o.tmask == 0) {   // tile-uniform factor (scalar registers)
        reg_phase_apply<T, A>(a, (T)o.m[0], (T)o.m[1], real, o.rmask, o.rval);
        return;
    }
    const bool lane_ok = ((uint32_t)t & o.tmask) == o.tval;
    reg_phase_apply<T, A>(a, lane_ok ? (T)o.m[0] : T(1), lane_ok ? (T)o.m[1] : T(0), real, o.rmask, o.rval);
}

// Moves the tile into another bit assignment through LDS. The maps hold
// swizzled tile indices (the swizzle is linear, so thread and register parts
// combine with XOR).
template <class T, int A>
__device__ __forceinline__ void reg_remap(typename Vec<T>::type (&a)[A], typename Vec<T>::type *lds,
                                          const PhaseMap &from, const PhaseMap &to, int t) {
    __syncthreads();   // the previous users of the LDS tile are done
    const uint32_t tf = (uint32_t)from.thr[0][t & 31] ^ (uint32_t)from.thr[1][(t >> 5) & 31];
#pragma unroll
    for (int k = 0; k < A; ++k) lds[tf ^ (uint32_t)from.reg[k]] = a[k];
    __syncthreads();
    const uint32_t tt = (uint32_t)to.thr[0][t & 31] ^ (uint32_t)to.thr[1][(t >> 5) & 31];
#pragma unroll
    for (int k = 0; k < A; ++k) a[k] = lds[tt ^ (uint32_t)to.reg[k]];
}

template <class T, int A>
__global__ void __launch_bounds__(reg_kernel_max_threads(A, sizeof(typename Vec<T>::type))) k_pass_reg(typename Vec<T>::type *__restrict__ state,
                                                   const RegOp *__restrict__ ops, int n_ops,
                                                   const PhaseMap *__restrict__ maps, const PassMem *__restrict__ pm,
                                                   uint64_t nonlocal_mask, int64_t n_tiles, int r) {
    using V = typename Vec<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    V *lds = reinterpret_cast<V *>(smem);
    const int t = threadIdx.x;
    (void)r;
    const uint64_t lt = pm->lthr[0][t & 31] | pm->lthr[1][(t >> 5) & 31];
    const uint64_t st = pm->sthr[0][t & 31] | pm->sthr[1][(t >> 5) & 31];
    for (int64_t tid = blockIdx.x; tid < n_tiles; tid += gridDim.x) {
        const uint64_t base = pdep64((uint64_t)tid, nonlocal_mask);
        V a[A];
#pragma unroll
        for (int k = 0; k < A; ++k) a[k] = load_stream<T>(&state[base | lt | pm->lreg[k]]);
        int cur = pm->first_map;
        for (int i = 0; i < n_ops; ++i) {
            const RegOp &o = ops[i];
            if (o.kind == R_REMAP) {
                reg_remap<T, A>(a, lds, maps[cur], maps[o.rbit], t);
                cur = o.rbit;
                continue;
            }
            if ((base & o.gmask) != o.gval) continue;   // tile-uniform
            if (o.kind == R_PHASE || o.kind == R_PHASE_REAL) reg_phase<T, A>(a, o, t);
            else reg_gate<T, A, false>(a, o, t);
        }
#pragma unroll
        for (int k = 0; k < A; ++k) store_stream<T>(&state[base | st | pm->sreg[k]], a[k]);
    }
}

// Batched register kernel: one launch runs one pass of each of many states
// (of any qubit count that shares the workgroup shape). Workgroup b belongs to
// the item whose tile range [tile_begin, next tile_begin) holds b; the items
// are sorted by tile_begin, so a binary search on tile-uniform values finds it.
// When the item asks for a readout, each tile also reduces its output (FP64
// sum of |a|^2, largest |a|^2 with the lowest index on ties) into
// partials[tile]; k_readout_final combines the tiles of a state.
template <class T, int A>
__global__ void __launch_bounds__(reg_kernel_max_threads(A, sizeof(typename Vec<T>::type))) k_batch_reg(const BatchItem *__restrict__ items,
                                                                          int n_items,
                                                                          const char *__restrict__ progs,
                                                                          char *__restrict__ states,
                                                                          char *__restrict__ parts) {
    using V = typename Vec<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red_s[16], red_m[16];
    __shared__ uint64_t red_i[16];
    V *lds = reinterpret_cast<V *>(smem);
    const int t = threadIdx.x;
    const int64_t b = blockIdx.x;
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].tile_begin <= b) lo = mid; else hi = mid - 1;
    }
    const BatchItem &it = items[lo];
    V *__restrict__ state = reinterpret_cast<V *>(states + it.state_off);
    const RegOp *__restrict__ ops = reinterpret_cast<const RegOp *>(progs + it.ops_off);
    const PhaseMap *__restrict__ maps = reinterpret_cast<const PhaseMap *>(progs + it.maps_off);
    const PassMem *__restrict__ pm = reinterpret_cast<const PassMem *>(progs + it.pm_off);
    const int n_ops = it.n_ops;
    const int64_t tid = b - it.tile_begin;
    const uint64_t base = pdep64((uint64_t)tid, it.nonlocal_mask);
    const uint64_t lt = pm->lthr[0][t & 31] | pm->lthr[1][(t >> 5) & 31];
    const uint64_t st = pm->sthr[0][t & 31] | pm->sthr[1][(t >> 5) & 31];
    V a[A];
#pragma unroll
    for (int k = 0; k < A; ++k) a[k] = state[base | lt | pm->lreg[k]];
    int cur = pm->first_map;
    for (int i = 0; i < n_ops; ++i) {
        const RegOp &o = ops[i];
        if (o.kind == R_REMAP) {
            reg_remap<T, A>(a, lds, maps[cur], maps[o.rbit], t);
            cur = o.rbit;
            continue;
        }
        if ((base & o.gmask) != o.gval) continue;
        if (o.kind == R_PHASE || o.kind == R_PHASE_REAL) reg_phase<T, A>(a, o, t);
        else reg_gate<T, A, true>(a, o, t);
    }
#pragma unroll
    for (int k = 0; k < A; ++k) state[base | st | pm->sreg[k]] = a[k];
    if (it.partials_off < 0) return;
    double sum = 0.0, best = -1.0;
    uint64_t bi = ~uint64_t(0);
#pragma unroll
    for (int k = 0; k < A; ++k) {
        const double p = (double)a[k].x * (double)a[k].x + (double)a[k].y * (double)a[k].y;
        const uint64_t idx = base | st | pm->sreg[k];
        sum += p;
        if (p > best || (p == best && idx < bi)) { best = p; bi = idx; }
    }
    // lanes of a wave first (only lanes that exist: width = min(64, blockDim)), then waves in order
    const int width = blockDim.x < 64 ? (int)blockDim.x : 64;
    for (int off = width >> 1; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, width);
        const double ob = __shfl_xor(best, off, width);
        const uint64_t oi = __shfl_xor(bi, off, width);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if ((t & 63) == 0) { red_s[t >> 6] = sum; red_m[t >> 6] = best; red_i[t >> 6] = bi; }
    __syncthreads();
    if (t == 0) {
        const int waves = (blockDim.x + 63) >> 6;
        for (int w = 1; w < waves; ++w) {
            sum += red_s[w];
            if (red_m[w] > best || (red_m[w] == best && red_i[w] < bi)) { best = red_m[w]; bi = red_i[w]; }
        }
        ReadoutPartial r;
        r.sum = sum;
        r.pmax = best;
        r.idx = bi;
        r.pad = 0;
        reinterpret_cast<ReadoutPartial *>(parts + it.partials_off)[tid] = r;
    }
}

// One workgroup per task: combines a state's per-tile readout partials in
// tile order.
__global__ void __launch_bounds__(256) k_readout_final(const ReadoutTask *__restrict__ tasks) {
    __shared__ double ss[256], sm[256];
    __shared__ uint64_t si[256];
    const ReadoutTask &tk = tasks[blockIdx.x];
    double sum = 0.0, best = -1.0;
    uint64_t bi = ~uint64_t(0);
    for (int64_t i = threadIdx.x; i < tk.n; i += 256) {
        const ReadoutPartial p = tk.partials[i];
        sum += p.sum;
        if (p.pmax > best || (p.pmax == best && p.idx < bi)) { best = p.pmax; bi = p.idx; }
    }
    ss[threadIdx.x] = sum;
    sm[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            const double o = sm[threadIdx.x + w];
            const uint64_t oi = si[threadIdx.x + w];
            if (o > sm[threadIdx.x] || (o == sm[threadIdx.x] && oi < si[threadIdx.x])) { sm[threadIdx.x] = o; si[threadIdx.x] = oi; }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ReadoutPartial r;
        r.sum = ss[0];
        r.pmax = sm[0];
        r.idx = si[0];
        r.pad = 0;
        *tk.out = r;
    }
}

// Plain streaming copy, 16 bytes per lane per iteration.
__global__ void __launch_bounds__(256) k_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
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

// Per-block partials of the comparison of two states of any amplitude types:
// Re and Im of sum conj(a)*b, sum |a|^2, sum |b|^2 (FP64), and max |a - b|.
template <class TA, class TB>
__global__ void __launch_bounds__(256) k_compare(const typename Vec<TA>::type *a, const typename Vec<TB>::type *b,
                                                  uint64_t n, double *partial) {
    __shared__ double sh[5][256];
    double re = 0, im = 0, na = 0, nb = 0, mx = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        re += ar * br + ai * bi;
        im += ar * bi - ai * br;
        na += ar * ar + ai * ai;
        nb += br * br + bi * bi;
        const double dr = ar - br, di = ai - bi;
        mx = fmax(mx, sqrt(dr * dr + di * di));
    }
    const int t = threadIdx.x;
    sh[0][t] = re; sh[1][t] = im; sh[2][t] = na; sh[3][t] = nb; sh[4][t] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + w];
            sh[4][t] = fmax(sh[4][t], sh[4][t + w]);
        }
        __syncthreads();
    }
    if (t < 5) partial[5 * blockIdx.x + t] = sh[t][0];
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

uint64_t max_blocks(int block) { return 0xffffffffull / (uint64_t)block; }

int grid_for(uint64_t n, int block) {
    uint64_t g = (n + block - 1) / block;
    if (g > (uint64_t(1) << 24)) g = uint64_t(1) << 24;
    if (g > max_blocks(block)) g = max_blocks(block);
    if (g < 1) g = 1;
    return (int)g;
}

int64_t tile_grid(int64_t n_tiles, int block) {
    int64_t g = n_tiles < (int64_t(1) << 22) ? n_tiles : (int64_t(1) << 22);
    if ((uint64_t)g > max_blocks(block)) g = (int64_t)max_blocks(block);
    return g;
}

// The per-pass kernels' dynamic LDS limit, raised once per kernel (before the
// first launch, and before any stream capture that launches it).
template <class T>
hipError_t pass_lds_attr() {
    static const hipError_t e =
        hipFuncSetAttribute((const void *)k_pass_lds<T>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    return e;
}
template <class T, int A>
hipError_t pass_reg_attr() {
    static const hipError_t e =
        hipFuncSetAttribute((const void *)k_pass_reg<T, A>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    return e;
}
template <class T>
hipError_t pass_attrs() {
    for (hipError_t e : {pass_lds_attr<T>(), pass_reg_attr<T, 2>(), pass_reg_attr<T, 4>(), pass_reg_attr<T, 8>(),
                         pass_reg_attr<T, 16>(), pass_reg_attr<T, 32>()})
        if (e != hipSuccess) return e;
    return hipSuccess;
}

template <class T>
hipError_t pass_lds(void *state, const DevOp *ops, int n_ops, const PassArgs &a, hipStream_t s) {
    pass_lds_attr<T>();
    const size_t lds = 3 * 64 * sizeof(uint64_t) + (size_t(1) << a.r) * sizeof(typename Vec<T>::type);
    const int64_t grid = tile_grid(a.n_tiles, a.B);
    hipLaunchKernelGGL(k_pass_lds<T>, dim3((unsigned)grid), dim3(a.B), lds, s,
                       reinterpret_cast<typename Vec<T>::type *>(state), ops, n_ops, a);
    return hipGetLastError();
}

template <class T, int A>
hipError_t pass_reg(void *state, const RegOp *ops, int n_ops, const PhaseMap *maps, const PassMem *pm,
                    uint64_t nonlocal, int64_t n_tiles, int r, int B, size_t lds, hipStream_t s) {
    pass_reg_attr<T, A>();
    const int64_t grid = tile_grid(n_tiles, B);
    hipLaunchKernelGGL((k_pass_reg<T, A>), dim3((unsigned)grid), dim3(B), lds, s,
                       reinterpret_cast<typename Vec<T>::type *>(state), ops, n_ops, maps, pm, nonlocal, n_tiles, r);
    return hipGetLastError();
}

template <class T>
hipError_t pass_reg_any(int A, void *state, const RegOp *ops, int n_ops, const PhaseMap *maps, const PassMem *pm,
                        uint64_t nonlocal, int64_t n_tiles, int r, int B, size_t lds, hipStream_t s) {
    switch (A) {
        case 2: return pass_reg<T, 2>(state, ops, n_ops, maps, pm, nonlocal, n_tiles, r, B, lds, s);
        case 4: return pass_reg<T, 4>(state, ops, n_ops, maps, pm, nonlocal, n_tiles, r, B, lds, s);
        case 8: return pass_reg<T, 8>(state, ops, n_ops, maps, pm, nonlocal, n_tiles, r, B, lds, s);
        case 16: return pass_reg<T, 16>(state, ops, n_ops, maps, pm, nonlocal, n_tiles, r, B, lds, s);
        case 32: return pass_reg<T, 32>(state, ops, n_ops, maps, pm, nonlocal, n_tiles, r, B, lds, s);
        default: return hipErrorInvalidValue;
    }
}

template <class T, int A>
hipError_t batch_reg(const BatchItem *items, int n_items, int64_t total_tiles, int B, size_t lds, char *arena,
                     hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipFuncSetAttribute((const void *)k_batch_reg<T, A>, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL((k_batch_reg<T, A>), dim3((unsigned)total_tiles), dim3(B), lds, s, items, n_items,
                       (const char *)arena, arena, arena);
    return hipGetLastError();
}

template <class T>
hipError_t batch_reg_any(int A, const BatchItem *items, int n_items, int64_t total_tiles, int B, size_t lds,
                         char *arena, hipStream_t s) {
    switch (A) {
        case 2: return batch_reg<T, 2>(items, n_items, total_tiles, B, lds, arena, s);
        case 4: return batch_reg<T, 4>(items, n_items, total_tiles, B, lds, arena, s);
        case 8: return batch_reg<T, 8>(items, n_items, total_tiles, B, lds, arena, s);
        case 16: return batch_reg<T, 16>(items, n_items, total_tiles, B, lds, arena, s);
        case 32: return batch_reg<T, 32>(items, n_items, total_tiles, B, lds, arena, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_pass_reg(int dtype, int A, void *state, const RegOp *ops, int n_ops, const PhaseMap *maps,
                           const PassMem *pm, uint64_t nonlocal, int64_t n_tiles, int r, int B, size_t lds_bytes,
                           hipStream_t s) {
    return dtype == HGM_C64 ? pass_reg_any<float>(A, state, ops, n_ops, maps, pm, nonlocal, n_tiles, r, B, lds_bytes, s)
                            : pass_reg_any<double>(A, state, ops, n_ops, maps, pm, nonlocal, n_tiles, r, B, lds_bytes, s);
}

hipError_t launch_pass_lds(int dtype, void *state, const DevOp *ops, int n_ops, const PassArgs &a, hipStream_t s) {
    return dtype == HGM_C64 ? pass_lds<float>(state, ops, n_ops, a, s) : pass_lds<double>(state, ops, n_ops, a, s);
}

hipError_t prepare_pass_kernels(int dtype) { return dtype == HGM_C64 ? pass_attrs<float>() : pass_attrs<double>(); }

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    // one 16-byte element per thread over the whole buffer: a flat grid keeps
    // far more requests in flight than a grid-stride loop (6.2 vs 4.9 TB/s)
    const uint64_t n16 = bytes / 16;
    const int blk = 256;
    uint64_t g = (n16 + blk - 1) / blk;
    if (g > max_blocks(blk)) g = max_blocks(blk);
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)g), dim3(blk), 0, s, (uint4 *)dst, (const uint4 *)src, n16);
    return hipGetLastError();
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

hipError_t launch_compare(int dtype_a, const void *a, int dtype_b, const void *b, uint64_t n, double *partial,
                          int blocks, hipStream_t s) {
    const bool fa = dtype_a == HGM_C64, fb = dtype_b == HGM_C64;
    if (fa && fb)
        hipLaunchKernelGGL((k_compare<float, float>), dim3(blocks), dim3(256), 0, s, (const float2 *)a, (const float2 *)b, n, partial);
    else if (fa)
        hipLaunchKernelGGL((k_compare<float, double>), dim3(blocks), dim3(256), 0, s, (const float2 *)a, (const double2 *)b, n, partial);
    else if (fb)
        hipLaunchKernelGGL((k_compare<double, float>), dim3(blocks), dim3(256), 0, s, (const double2 *)a, (const float2 *)b, n, partial);
    else
        hipLaunchKernelGGL((k_compare<double, double>), dim3(blocks), dim3(256), 0, s, (const double2 *)a, (const double2 *)b, n, partial);
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

hipError_t launch_batch_reg(int dtype, int A, int B, const BatchItem *items, int n_items, int64_t total_tiles,
                            size_t lds_bytes, void *arena, hipStream_t s) {
    if (n_items < 1 || total_tiles < 1 || B < 1 || (uint64_t)total_tiles > max_blocks(B) || B > reg_kernel_max_threads(A, amp_bytes(dtype)) ||
        !arena)
        return hipErrorInvalidValue;
    char *base = static_cast<char *>(arena);
    return dtype == HGM_C64 ? batch_reg_any<float>(A, items, n_items, total_tiles, B, lds_bytes, base, s)
                            : batch_reg_any<double>(A, items, n_items, total_tiles, B, lds_bytes, base, s);
}

hipError_t launch_readout_final(const ReadoutTask *tasks, int n_tasks, hipStream_t s) {
    if (n_tasks < 1) return hipSuccess;
    hipLaunchKernelGGL(k_readout_final, dim3((unsigned)n_tasks), dim3(256), 0, s, tasks);
    return hipGetLastError();
}

}  // namespace hgm
