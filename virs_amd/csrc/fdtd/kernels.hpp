// Device kernels of the FDTD step for gfx950 (CDNA4, wave64).
//
// Semantics follow the reference GPU step (src/simulator/simulator_gpu.cu):
//   dense update      :38-103   -> k_dense
//   boundary biquad   :105-150  -> fused into k_dense (per-plane slices of the
//                                  compact list), or k_tail over the list
//   source, receiver  :59-63,214-218,292-300 -> k_dense (hard) / k_tail (points)
//   adjoint           (new)     -> k_dense<ADJ> + k_tail_adj
// Arithmetic is written in the reference's operand order and the file is
// compiled without FP contraction, so an FP32/FP64 run is bit-identical to a
// CPU evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#pragma clang fp contract(off)

namespace virs {

constexpr int kBlock = 256;   // 4 waves of 64
// LDS staging of the stencil's planes (k_dense LY): LY = 1 for rows of at
// most kBlock / 2 vectors (a block's two halo rows take one vector per
// thread), LY = 2 for rows of at most kBlock vectors (two per thread)
constexpr int kHaloVecs = kBlock / 2;
constexpr int kWaves = kBlock / 64;
constexpr int kSlots = 64;    // spread of the per-step |p| maximum atomics
constexpr int kMaxSponge = 64;
constexpr int kMaxPoints = 64;

// Device-resident step state.  nring[k] is the step index that the launches
// of rotation phase k compute; each step's last launch writes the next
// phase's entry, which no block of the current launch reads, so the counter
// advances without atomics and a captured graph of whole rotation cycles
// replays correctly.  exp/maxabs are rings indexed by step & 3 holding each
// stored field's power-of-two scale exponent and the largest |value| the
// stencil launch computed for it (physical units, float bits: non-negative
// floats order like their bit patterns).  The boundary correction, a bounded
// damping term, and the point sources, which enter the exponent's estimate
// themselves, do not contribute to it.
// The RMS early stop on receiver 0 (simulator_gpu.cu:303-330) is evaluated
// on the device at the end of every step.  The step that fires records
// itself in stop_at and hands the next step a counter with kStopBit set;
// every launch that reads a counter with the bit exits at once (the list
// launch still passing the counter on), so batches and graph replays stop
// without a host round trip and without a load beyond the counter itself.
constexpr int kRmsWindow = 100, kRmsMinStep = 3000;
constexpr long long kStopBit = 1ll << 62;
struct StepCtl {
    long long nring[3];
    int err;
    int pad;
    int exp[4];
    unsigned int maxabs[4][kSlots];
    int es_enable;
    int pad1;
    long long stop_at;   // LLONG_MAX until the rule fires
    float es_target, rms_state, percent;
    int pad2;
};

// 16-bit float storage (IEEE binary16) and bfloat16 storage.
typedef _Float16 f16;
struct bf16 {
    uint16_t bits;
};

template <typename T>
struct Store;
template <>
struct Store<float> {
    using C = float;
    static constexpr bool kScaled = false;
    __device__ static float ld(float v) { return v; }
    __device__ static float st(float c) { return c; }
};
template <>
struct Store<double> {
    using C = double;
    static constexpr bool kScaled = false;
    __device__ static double ld(double v) { return v; }
    __device__ static double st(double c) { return c; }
};
template <>
struct Store<f16> {
    using C = float;
    static constexpr bool kScaled = true;
    __device__ static float ld(f16 v) { return (float)v; }
    __device__ static f16 st(float c) { return (f16)c; } // round to nearest even
};
template <>
struct Store<bf16> {
    using C = float;
    static constexpr bool kScaled = true;
    __device__ static float ld(bf16 v) { return __uint_as_float((uint32_t)v.bits << 16); }
    __device__ static bf16 st(float c)
    {
        uint32_t u = __float_as_uint(c);
        bf16 r;
        if ((u & 0x7fffffffu) > 0x7f800000u) {
            r.bits = (uint16_t)((u >> 16) | 0x40u); // quiet NaN
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);        // round to nearest even
            r.bits = (uint16_t)(u >> 16);
        }
        return r;
    }
};

// Stochastic rounding to a 16-bit format.  A counter-based hash of
// (cell pair, step, salt) gives 16 random bits per cell (one 32-bit hash
// serves the cells 2k and 2k+1: low half, high half).  They are masked to
// the bits below the format's last place (13 for FP16, 16 for BF16), added
// to the float's bit pattern, and the result converted toward zero, which
// drops those low bits.  In the normal range that is exactly stochastic
// rounding: E[stored] = value, so increments smaller than half an ulp are
// not lost step after step; in FP16's subnormal range the conversion
// truncates.  The mixing multiply is the full-rate 24-bit one (the xor-shift
// before it folds the top byte into the low bits).
constexpr uint32_t kHashK1 = 0x9E3779B1u, kHashK2 = 0x1B3C6Du, kHashK3 = 0x85EBCA77u;
__device__ inline uint32_t hash_pair(uint32_t lo_k1, uint32_t hk)
{
    uint32_t h = lo_k1 + hk;
    h ^= h >> 16;
    h = __umul24(h, kHashK2);
    h ^= h >> 15;
    return h;
}
// hk term of the hash for cells idx (any of a pair) at step n
__device__ inline uint32_t hash_hk(unsigned long long idx, long long n, uint32_t salt)
{
    return ((uint32_t)(idx >> 32) + (uint32_t)n) * kHashK3 + salt;
}
__device__ inline uint32_t mix16(unsigned long long idx, long long n, uint32_t salt)
{
    const unsigned long long pair = idx & ~1ull;
    const uint32_t h = hash_pair((uint32_t)pair * kHashK1, hash_hk(pair, n, salt));
    return (idx & 1ull) ? (h >> 16) : (h & 0xffffu);
}

constexpr uint32_t kSaltDense = 0u, kSaltBoundary = 0x68E31DA4u, kSaltSource = 0xB5297A4Du;

// x with the random bits r16 added below the 16-bit format's last place;
// converting the result toward zero (cvt2_truncated) rounds x stochastically.
template <typename T>
constexpr uint32_t sr_mask()
{
    return __is_same(T, bf16) ? 0xffffu : 0x1fffu;
}
template <typename T>
__device__ inline float sr_bits(float x, uint32_t r16)
{
    return __uint_as_float(__float_as_uint(x) + (r16 & sr_mask<T>()));
}

// Converts two floats to T rounding toward zero.
template <typename T>
__device__ inline void cvt2_truncated(float a, float b, T &ta, T &tb)
{
    if constexpr (__is_same(T, bf16)) {
        ta.bits = (uint16_t)(__float_as_uint(a) >> 16);
        tb.bits = (uint16_t)(__float_as_uint(b) >> 16);
    } else {
        const auto h = __builtin_amdgcn_cvt_pkrtz(a, b);
        ta = (T)h[0];
        tb = (T)h[1];
    }
}

// Single-value storage rounding (used off the hot loop): nearest-even, or
// stochastic for the 16-bit formats when enabled.
template <typename T>
__device__ inline T store_round(typename Store<T>::C x, bool sr, unsigned long long idx, long long n,
                                uint32_t salt)
{
    if constexpr (Store<T>::kScaled) {
        if (sr) {
            T a, b;
            cvt2_truncated<T>(sr_bits<T>(x, mix16(idx, n, salt)), 0.f, a, b);
            return a;
        }
    }
    return Store<T>::st(x);
}

template <typename T, int V>
struct alignas(sizeof(T) * V) Vec {
    T v[V];
};

// Raw vector memory access of B bytes, optionally non-temporal (streamed
// data: read or written once per step, kept out of the way of the reused
// p^n planes in L2).
template <int B>
struct Raw;
template <>
struct Raw<16> {
    typedef unsigned int type __attribute__((ext_vector_type(4)));
};
template <>
struct Raw<8> {
    typedef unsigned int type __attribute__((ext_vector_type(2)));
};
template <>
struct Raw<4> {
    typedef unsigned int type;
};
template <>
struct Raw<2> {
    typedef unsigned short type;
};
template <>
struct Raw<1> {
    typedef unsigned char type;
};

template <typename VT, bool NT>
__device__ inline VT vload(const VT *p)
{
    using R = typename Raw<sizeof(VT)>::type;
    R r;
    if constexpr (NT) r = __builtin_nontemporal_load(reinterpret_cast<const R *>(p));
    else r = *reinterpret_cast<const R *>(p);
    VT v;
    __builtin_memcpy(&v, &r, sizeof(VT));
    return v;
}

template <typename VT, bool NT>
__device__ inline void vstore(VT *p, const VT &v)
{
    using R = typename Raw<sizeof(VT)>::type;
    R r;
    __builtin_memcpy(&r, &v, sizeof(VT));
    if constexpr (NT) __builtin_nontemporal_store(r, reinterpret_cast<R *>(p));
    else *reinterpret_cast<R *>(p) = r;
}

__device__ inline float exp2i(int e) { return __int_as_float((e + 127) << 23); }

// sum + h, h the low (HI false) or high FP16 half of w, as one
// v_fma_mix_f32 computing 1*h + sum: the product is exact, so the sum is
// rounded once, bit for bit the result of converting h and adding it
template <bool HI>
__device__ inline float add_f16_half(float sum, uint32_t w)
{
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, 1.0, %1, %2 op_sel:[0,1,0] op_sel_hi:[0,1,0]" : "=v"(r) : "v"(w), "v"(sum));
    else
        asm("v_fma_mix_f32 %0, 1.0, %1, %2 op_sel_hi:[0,1,0]" : "=v"(r) : "v"(w), "v"(sum));
    return r;
}

// a * h + c rounded once, h the FP16 half of w (high half if HI), with a
// and c FP32
template <bool HI>
__device__ inline float fma_f16_half(float a, uint32_t w, float c)
{
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,1,0] op_sel_hi:[0,1,0]" : "=v"(r) : "v"(a), "v"(w), "v"(c));
    else
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,1,0]" : "=v"(r) : "v"(a), "v"(w), "v"(c));
    return r;
}

// Orders a wave's LDS accesses around a lane-to-lane exchange inside the
// wave: the hardware performs one wave's LDS operations in issue order, so
// only the compiler's ordering needs fixing, which a wavefront-scope fence
// does without an instruction.
__device__ inline void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Divergence check of the list and point launches: a non-finite corrected
// boundary value or receiver sample marks the run (StepCtl::err), which
// virs_synchronize reports in every precision.
template <typename C>
__device__ inline void flag_nonfinite(StepCtl *ctl, C v)
{
    if (!__builtin_isfinite(v)) ctl->err = 1;
}

template <typename C>
__device__ inline C wave_max(C v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

template <typename C>
__device__ inline C wave_sum(C v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o);
    return v;
}

__device__ inline int wave_max_int(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// Cell code of FP32/FP64 fields, one byte per cell, built on the device from
// the reference flags (16-bit fields use k_build_code16's half-byte code):
//   bits 0-5  neighbour is a wall (flag 1): +x 1, -x 2, +y 4, -y 8, +z 16, -z 32
//   bits 6-7  0 update (flag 0), 3 update + boundary filter (interior flag 2),
//             1 write zero (flag 1 and the domain border), 2 hard source (flag 3)
// The stencil reads this one byte per cell instead of seven flags.
enum : uint8_t { kTypeUpdate = 0, kTypeZero = 1, kTypeHard = 2, kTypeBoundary = 3 };

__global__ void k_build_code(const uint8_t *__restrict__ flags, uint8_t *__restrict__ code, int nx,
                             int ny, int nz)
{
    const long long sxy = (long long)nx * ny, n = sxy * nz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i / sxy);
        const long long r = i - (long long)z * sxy;
        const int y = (int)(r / nx);
        const int x = (int)(r - (long long)y * nx);
        uint8_t c;
        if (x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1) {
            c = kTypeZero << 6;
        } else {
            const uint8_t f = flags[i];
            if (f == 1) {
                c = kTypeZero << 6;
            } else if (f == 3) {
                c = kTypeHard << 6;
            } else {
                c = f == 2 ? (uint8_t)(kTypeBoundary << 6) : (uint8_t)0;
                if (flags[i + 1] == 1) c |= 1;
                if (flags[i - 1] == 1) c |= 2;
                if (flags[i + nx] == 1) c |= 4;
                if (flags[i - nx] == 1) c |= 8;
                if (flags[i + sxy] == 1) c |= 16;
                if (flags[i - sxy] == 1) c |= 32;
            }
        }
        code[i] = c;
    }
}

// Cell code of the 16-bit formats, half a byte per cell: cell i sits in the
// low nibble of byte i/2 when i is even, in the high nibble when odd.  In
// 16-bit storage every wall and border cell holds zero, so a wall neighbour
// contributes nothing to the plain sum of the six neighbours and the
// reference's substitution V(wall) = pc becomes a smaller centre weight:
// lap = (sum of neighbours - kw pc) ih2 with kw = 6 - (wall neighbours).
// Nibble: kw (0..6) for cells that update, kNib16Zero for walls and border
// cells, kNib16Hard for hard-source cells (bit 3 marks both, bit 0 tells
// them apart).
constexpr uint32_t kNib16Zero = 8, kNib16Hard = 9;
__global__ void k_build_code16(const uint8_t *__restrict__ flags, uint8_t *__restrict__ code, int nx,
                               int ny, int nz)
{
    const long long sxy = (long long)nx * ny, n = sxy * nz, nbytes = (n + 1) / 2;
    for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < nbytes;
         b += (long long)gridDim.x * blockDim.x) {
        uint32_t pair = 0;
        for (int h = 0; h < 2; ++h) {
            const long long i = 2 * b + h;
            uint32_t c = kNib16Zero; // the pad nibble past the last cell
            if (i < n) {
                const int z = (int)(i / sxy);
                const long long r = i - (long long)z * sxy;
                const int y = (int)(r / nx);
                const int x = (int)(r - (long long)y * nx);
                if (x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1 || flags[i] == 1) {
                    c = kNib16Zero;
                } else if (flags[i] == 3) {
                    c = kNib16Hard;
                } else {
                    const int walls = (flags[i + 1] == 1) + (flags[i - 1] == 1) + (flags[i + nx] == 1) +
                                      (flags[i - nx] == 1) + (flags[i + sxy] == 1) + (flags[i - sxy] == 1);
                    c = (uint32_t)(6 - walls);
                }
            }
            pair |= c << (4 * h);
        }
        code[b] = (uint8_t)pair;
    }
}

__device__ inline uint32_t code16_at(const uint8_t *__restrict__ code, long long i)
{
    return ((uint32_t)code[i >> 1] >> ((uint32_t)(i & 1) * 4u)) & 0xFu;
}

// The nibble of cell i in bits 0..3 (the rest garbage), one streaming byte load.
__device__ inline uint32_t load_nibble1(const uint8_t *__restrict__ code, long long i)
{
    return (uint32_t)__builtin_nontemporal_load(code + (i >> 1)) >> ((uint32_t)(i & 1) * 4u);
}

// The nibbles of V >= 2 cells that start at byte `off` of `code` (an even
// cell), cell j in bits 4j .. 4j+3, with one streaming load of V/2 bytes.
template <int V>
__device__ inline uint32_t load_nibbles(const uint8_t *__restrict__ code, uint32_t off)
{
    const uint8_t *b = code + off;
    if constexpr (V == 8) return __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(b));
    else if constexpr (V == 4) return (uint32_t)__builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(b));
    else return (uint32_t)__builtin_nontemporal_load(b);
}

// Zeroes the cells a 16-bit code marks as walls or border.
template <typename T>
__global__ void k_zero_walls(T *__restrict__ f, const uint8_t *__restrict__ code, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        if (code16_at(code, i) == kNib16Zero) f[i] = Store<T>::st(0.f);
}

// All-ones if bit `bit` of c is set, else zero (one signed bit-field extract).
__device__ inline uint32_t bit_mask(uint32_t c, int bit) { return (uint32_t)__builtin_amdgcn_sbfe((int)c, bit, 1); }
// Bitwise select, a where m is set, else b, as one v_bfi_b32 (the compiler
// would otherwise rebuild a compare + conditional move from the mask).
__device__ inline uint32_t bfi(uint32_t m, uint32_t a, uint32_t b)
{
    uint32_t r;
    asm volatile("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}
__device__ inline float bsel(uint32_t m, float a, float b)
{
    return __uint_as_float(bfi(m, __float_as_uint(a), __float_as_uint(b)));
}
__device__ inline double bsel(uint32_t m, double a, double b)
{
    const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
    const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
    const unsigned long long lo = bfi(m, (uint32_t)ua, (uint32_t)ub);
    const unsigned long long hi = bfi(m, (uint32_t)(ua >> 32), (uint32_t)(ub >> 32));
    return __longlong_as_double((long long)(lo | (hi << 32)));
}

struct Sponge {
    int w;
    float a[kMaxSponge], b[kMaxSponge], g[kMaxSponge];
};

__device__ inline int edge_dist(int x, int y, int z, int nx, int ny, int nz)
{
    return min(min(min(x, nx - 1 - x), min(y, ny - 1 - y)), min(z, nz - 1 - z));
}

struct BoundaryState {
    long long nb;
    const uint8_t *bmat;       // material row per entry (zeros for one material)
    // Filter state.  Rings are indexed by step parity: on entry to step n,
    // slot n&1 holds the value of step n-2 and slot (n+1)&1 that of step
    // n-1; step n overwrites slot n&1, so a step writes one value per ring
    // instead of shifting the state.
    // FP32/FP64 fields: ws = the direct-form-II state w as an FP64 ring, 8
    // bytes read twice and written once per entry and step.  xs/ys = the
    // history of the filter inputs d (in the compute type C, which holds d
    // exactly: a difference of two C values) and outputs y (FP64), rings
    // written only when non-null (the recording forward's last two steps
    // before each snapshot, which seed the adjoint's trace) and never read
    // by a step.
    // 16-bit fields: ts = the transposed-direct-form-II states s1, s2,
    // computed in FP64 and stored in FP32, read and rewritten in place every
    // step (16 bytes per entry and step).
    // Reverse: xs = dbar, ys = ybar, both FP64.
    void *ws[2];
    void *ts[2];
    void *xs[2];
    void *ys[2];
    const double *coeffs;      // [nmat][5] b0 b1 b2 a1 a2
    int nmat;
    float kconst;              // (c*dt)/h*0.95f
    // forward boundary trace for the adjoint, in the compute type C of the
    // field (float for FP32, double for FP64): d^n and y^n of every entry,
    // [slots][nb] with slot = n - trace_base + 2; null = off
    void *trace_d, *trace_y;
    long long trace_base, trace_slots;
};

struct DenseParams {
    const void *P;   // p^n      (forward) / adjoint of p^{n+1} (reverse)
    const void *Q;   // p^{n-1}  (forward) / adjoint of p^{n+2} (reverse)
    void *R;         // p^{n+1}  (forward) / adjoint of p^n     (reverse)
    const uint8_t *code;
    int nx, ny, nz;
    long long sxy;
    long long nvec_plane; // vectors of V cells per z-plane
    int nchunks_plane;    // 256-thread chunks per z-plane
    int zranges;          // > 0: tile mode, block = (chunk, one of zranges z-ranges)
    float c2dt2, ih2;
    StepCtl *ctl;
    int slot;             // rotation phase: reads ctl->nring[slot]
    const double *hard_sig;
    long long hard_len;
    int hard_ref;         // 1: the reference pulse (10 for n < 10)
    int nsoft;            // soft sources: their |signal| enters the scale estimate
    const double *soft_sig;
    long long soft_len;
    Sponge sp;
    int sr;               // stochastic storage rounding (16-bit formats)
    // Boundary filter inside the stencil pass: the list entries whose cells
    // lie in the 64*V cells of wave w of chunk ch at plane z are the
    // contiguous slice wtab[(ch*kWaves + w)*nz + z] = {first entry, count}
    // of the ascending list (rows z-contiguous, so a wave's consecutive
    // planes share a scalar cache line).  Null: the list runs in k_tail.
    const uint2 *wtab;
    const uint32_t *bidx;
    BoundaryState bs;
};

// Biquad step and boundary correction for one list entry (simulator_gpu.cu:126-148).
// The reference's filter is y = b0 d + b1 d^{n-1} + b2 d^{n-2} - a1 y^{n-1} - a2 y^{n-2}
// (direct form I) with float state.  The same filter runs in double here, in
// a form that keeps two values of state per entry instead of four:
//   direct form II (FP32/FP64 fields), w1 = w^{n-1}, w2 = w^{n-2} in FP64:
//     w = d - a1 w1 - a2 w2,   y = b0 w + b1 w1 + b2 w2
//   transposed direct form II (16-bit fields), states s1, s2 stored in FP32:
//     y = b0 d + s1,   s1' = b1 d - a1 y + s2,   s2' = b2 d - a2 y
// Every form gives the same y in exact arithmetic.  The transposed form's
// states are of the size of y, so their FP32 rounding (~6e-8 relative) stays
// far below the 16-bit fields' storage noise; form II's w can exceed d by
// the filter's low-frequency gain (1e4 for the reference's default
// high-pass row), so it is kept in FP64.  Then
//   R = R - K (0.1 y + 0.9 d) 0.4f             (evaluated in double)
// The arithmetic shared by the list launch and the stencil's boundary phase
// (one expression order for both).
__device__ inline double df2_state(const double *cf, double in, double w1, double w2)
{
    return in - cf[3] * w1 - cf[4] * w2;
}
__device__ inline double df2_out(const double *cf, double w, double w1, double w2)
{
    return cf[0] * w + cf[1] * w1 + cf[2] * w2;
}
__device__ inline double tdf2_out(const double *cf, double in, double s1) { return cf[0] * in + s1; }
__device__ inline double tdf2_s1(const double *cf, double in, double y, double s2)
{
    return cf[1] * in - cf[3] * y + s2;
}
__device__ inline double tdf2_s2(const double *cf, double in, double y) { return cf[2] * in - cf[4] * y; }
template <typename C>
__device__ inline C corrected(C r, float kconst, double yv, double in)
{
    return (C)((double)r - (double)kconst * (yv * 0.1 + in * 0.9) * (double)0.4f);
}

// Stores of entry e at step n with 16-bit fields: the new states.
__device__ inline void store_tdf2(const BoundaryState &bs, long long e, double s1, double s2)
{
    ((float *)bs.ts[0])[e] = (float)s1;
    ((float *)bs.ts[1])[e] = (float)s2;
}
// Stores of entry e at step n with FP32/FP64 fields: the new state w, the
// d/y history when its rings are given and, when recording, the trace of
// the adjoint's segment.
template <typename C>
__device__ inline void store_df2(const BoundaryState &bs, long long e, long long n, C dif, double yv, double wv)
{
    ((double *)bs.ws[n & 1])[e] = wv;
    if (bs.xs[0]) {
        ((C *)bs.xs[n & 1])[e] = dif;
        ((double *)bs.ys[n & 1])[e] = yv;
    }
    if (bs.trace_d) {
        const long long slot = n - bs.trace_base + 2;
        if (slot >= 0 && slot < bs.trace_slots) {
            ((C *)bs.trace_d)[slot * bs.nb + e] = (C)(double)dif;
            ((C *)bs.trace_y)[slot * bs.nb + e] = (C)yv;
        }
    }
}

// Leapfrog 7-point update, simulator_gpu.cu:38-103 with a graded sponge:
//   lap = (V(x+1)+V(x-1)+V(y+1)+V(y-1)+V(z+1)+V(z-1) - 6 pc) * ih2
//   R   = (a pc - b Q) + g (c2dt2 lap),   V(nbr) = pc where nbr is a wall
// A thread owns V consecutive x cells of a row of the flattened z-plane and
// marches up in z: z neighbours stay in registers, x neighbours come from
// the vector and the adjacent lanes, y neighbours are the rows +-Nx of the
// same plane.  The next plane's loads are issued before the current plane
// is computed.  Every cell's update is computed and the result selected by
// cell type, without branches.
// LY (rows of at most LY * kHaloVecs vectors): each plane is staged in
// LDS.  The block writes its chunk of P(z), which its threads hold anyway,
// and one row on either side (LY halo vectors per thread, loaded with the
// plane), then reads the rows +-Nx and the end lanes' x neighbours back
// after a barrier.
// Two buffers alternate by plane, so one barrier per plane orders both the
// writes before the reads and the reads before the next overwrite.
// Otherwise each thread loads its rows +-Nx itself (L2 hits: in tile mode
// the neighbouring chunks march the same z together).
// 16-bit storage: arithmetic runs on the stored values, p^n in its units
// (value * 2^eP) and q and the result in p^{n+1}'s (2^eR, folded into the
// update's factors); power-of-two scaling commutes exactly with every
// rounding step, so the results equal a computation in physical units bit
// for bit.  Walls hold
// zero and the code carries the centre weight kw (k_build_code16); SR
// selects stochastic storage rounding.
// FUSED: the boundary biquad of the list entries runs inside the march (see
// the boundary phase in the plane body).  ADJ: the reverse sweep applies the same operator to the adjoint
// field; hard-source cells hold prescribed values, so their adjoint is zero.
// HARD (16-bit storage): the room has flag-3 cells.  Without them a wall or
// border cell (P = Q = 0 there) gets the update with a zero Laplacian
// factor, which is exactly +0, in place of two selects per cell.
template <typename T, int V, bool ADJ, bool FUSED, bool SPONGE, bool SR, bool HARD = true, int LY = 0>
#ifndef VIRS_DENSE_WAVES
#define VIRS_DENSE_WAVES 1 // minimum waves per SIMD the register allocation must allow
#endif
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(VIRS_DENSE_WAVES))) void k_dense(const DenseParams p)
{
    using S = Store<T>;
    using C = typename S::C;
    using VT = Vec<T, V>;
    using VB = Vec<uint8_t, V>;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    // XCD-aware order: the hardware places block b on XCD b % 8; give each
    // XCD a contiguous run of logical tiles so neighbouring chunks share an L2.
    const unsigned nbk = gridDim.x, bid = blockIdx.x;
    const unsigned logical = (nbk % 8u == 0u) ? (bid % 8u) * (nbk / 8u) + bid / 8u : bid;

    const long long n = p.ctl->nring[p.slot];
    if (n & kStopBit) return;
    C hard = (C)0;
    if (!ADJ) {
        if (p.hard_ref) hard = n < 10 ? (C)10 : (C)0;
        else if (n >= 0 && n < p.hard_len) hard = (C)p.hard_sig[n];
    }

    // kq: Q -> R's stored units; kout: P units -> R's stored units; kphys: P
    // units -> physical; kst / kst_inv: physical -> R's stored units and back.
    // The 16-bit update takes P in its stored units and Q in R's, with kout
    // folded into its factors (2 kout, c2dt2 ih2 kout, the sponge's a kout),
    // so it produces R's units directly: each is a power-of-two scaling of
    // the same operation in physical units, which rounds alike.
    C kq = (C)1, kout = (C)1, kphys = (C)1, kst = (C)1, kst_inv = (C)1;
    C hard_u = hard;
    if constexpr (S::kScaled) {
        const int ip = (int)((n - 1) & 3), iq = (int)((n - 2) & 3), ir = (int)(n & 3);
        const float mp = wave_max(__uint_as_float(p.ctl->maxabs[ip][lane]));
        const float mq = wave_max(__uint_as_float(p.ctl->maxabs[iq][lane]));
        float est = 16.f * fmaxf(mp, mq) + fabsf((float)hard);
        for (int k = 0; k < p.nsoft; ++k)
            if (n >= 0 && n < p.soft_len) est += fabsf((float)p.soft_sig[(long long)k * p.soft_len + n]);
        int eR = 0;
        if (est > 0.f && est < 3.0e38f) {
            int k2;
            frexpf(est, &k2); // est < 2^k2
            eR = min(max(14 - k2, -60), 60);
        }
        const int eP = p.ctl->exp[ip], eQ = p.ctl->exp[iq];
        kq = exp2i(eR - eQ);
        kout = exp2i(eR - eP);
        kphys = exp2i(-eP);
        kst = exp2i(eR);
        kst_inv = exp2i(-eR);
        hard_u = hard * exp2i(eR);
        if (bid == 0) {
            if (threadIdx.x == 0) {
                p.ctl->exp[ir] = eR;
                if (!(est < 3.0e38f)) p.ctl->err = 1;
            }
            if (threadIdx.x < kSlots) p.ctl->maxabs[(n + 1) & 3][threadIdx.x] = 0u;
        }
    }

    const T *__restrict__ P = (const T *)p.P;
    const T *__restrict__ Q = (const T *)p.Q;
    T *__restrict__ R = (T *)p.R;
    const long long sxy = p.sxy;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const C c2dt2 = (C)p.c2dt2, ih2 = (C)p.ih2;
    const C c2ih2 = (c2dt2 * ih2) * kout; // 16-bit scheme: c2dt2*ih2 folded into one factor, in R's units
    const C two_k = (C)2 * kout;
    const C nkq = -kq;

    __shared__ float sp_tab[SPONGE ? 3 : 1][SPONGE ? kMaxSponge : 1];
    if constexpr (SPONGE) {
        for (int k = threadIdx.x; k < p.sp.w; k += kBlock) {
            sp_tab[0][k] = S::kScaled ? p.sp.a[k] * (float)kout : p.sp.a[k];
            sp_tab[1][k] = p.sp.b[k];
            sp_tab[2][k] = p.sp.g[k];
        }
        __syncthreads();
    }
    float lmax = 0.f;

    // Work split over the (chunk, z) planes, flattened chunk-major.
    //  tile mode: block = chunk (logical % nchunks) x z-range (logical / nchunks),
    //    so the blocks running together are neighbouring chunks marching the
    //    same z and their y-halo rows meet in L2;
    //  otherwise: gridDim contiguous ranges of equal length.
    long long wbeg, wend;
    if (p.zranges > 0) {
        const long long ch = logical % (unsigned)p.nchunks_plane, zr = logical / (unsigned)p.nchunks_plane;
        const long long za = zr * nz / p.zranges, zb = (zr + 1) * nz / p.zranges;
        wbeg = ch * nz + za;
        wend = zr < p.zranges ? ch * nz + zb : wbeg;
    } else {
        const long long W = (long long)p.nchunks_plane * nz;
        const long long per = W / nbk, rem = W % nbk;
        wbeg = (long long)logical * per + min((long long)logical, rem);
        wend = wbeg + per + ((long long)logical < rem ? 1 : 0);
    }

    // cell codes of one plane: V bytes, or V nibbles in one word for 16-bit
    using CD = typename std::conditional<S::kScaled, uint32_t, VB>::type;
    struct Stage {
        VT pp, q, yn, ys; // P(z+1), Q(z), P(z,y+1), P(z,y-1) as stored
        VT h, h2;         // LY: this thread's halo vectors of P(z)
        CD cd;            // cell codes of plane z
        T tl, tr;         // P(z) left of lane 0 / right of lane 63
        // FUSED: the wave's list slice at plane z (first entry, count; wave
        // uniform) and lane L's entry first + L of it: cell, material,
        // filter-input and filter-output rings (lanes past the count hold
        // the slice's last entry)
        uint32_t efirst, ecount, eb, em;
        double ew0, ew1; // direct form II ring (FP32/FP64 fields)
        float es1, es2;  // transposed direct form II states (16-bit fields)
    };
    // FUSED: the cells' (d, stored p^{n+1}) of one wave and plane, by
    // position j in the lane's vector and lane, so an entry's lane can read
    // and correct its cell's value (wave-private, no barrier)
    struct alignas(2 * sizeof(C)) BSlot {
        C d, r;
    };
    __shared__ BSlot bslot[FUSED ? kWaves : 1][FUSED ? V : 1][64];
    // FUSED: the coefficient rows in LDS, so a lane's rows come without a
    // global load issued behind the next plane's (whose wait would drain
    // the prefetch)
    __shared__ double cf_lds[FUSED ? 5 * 255 : 1];
    if constexpr (FUSED) {
        for (int k = threadIdx.x; k < 5 * p.bs.nmat; k += kBlock) cf_lds[k] = p.bs.coeffs[k];
        __syncthreads();
    }

    // LY: two plane buffers of a chunk (slots 0..H-1 the row below it, H..
    // its own vectors, then the row above); the buffer alternates by plane
    static_assert(!(LY && FUSED), "the fused boundary keeps its own LDS use");
    static_assert(LY >= 0 && LY <= 2, "one or two halo vectors per thread");
    constexpr int kHalo = LY * kHaloVecs; // halo vectors per side at most
    using RawV = typename Raw<sizeof(VT)>::type;
    __shared__ RawV ybuf[LY ? 2 : 1][LY ? kHalo + kBlock + kHalo : 1];
    const int H = nx / V; // vectors per row; LY: at most kHalo
    unsigned pcount = 0;  // planes staged so far (LY)
    for (long long w = wbeg; w < wend;) {
        const int pchunk = (int)(w / nz);
        const int z0 = (int)(w - (long long)pchunk * nz);
        const int z1 = (int)min((long long)nz, (long long)z0 + (wend - w));
        w += z1 - z0;

        const long long v = (long long)pchunk * kBlock + threadIdx.x;
        const bool active = v < p.nvec_plane;
        // lanes past the plane's last vector read cell 0 of the plane, a
        // border cell whose code marks it zero, and store nothing
        const long long cell0 = active ? v * V : 0;
        const int y = (int)(cell0 / nx);
        const int x0 = (int)(cell0 - (long long)y * nx);
        // out-of-grid neighbours read a valid cell instead; only border
        // cells (written as zero) would use them
        const long long dyn = (active && y + 1 < ny) ? nx : 0;
        const long long dys = (active && y > 0) ? nx : 0;
        const bool need_l = active && lane == 0 && x0 > 0;
        const bool need_r = active && lane == 63 && x0 + V < nx;
        // per-lane byte offsets inside a plane (a plane is far below 4 GiB);
        // the plane's base address is uniform across the block
        const uint32_t oc = (uint32_t)cell0 * (uint32_t)sizeof(T);
        const uint32_t on = oc + (uint32_t)dyn * (uint32_t)sizeof(T);
        const uint32_t os = oc - (uint32_t)dys * (uint32_t)sizeof(T);
        const uint32_t ob = (uint32_t)cell0;
        // The offsets as used by the accesses.  An empty asm rewrites them
        // in place at every plane (same values), so their zero-extension
        // stays next to the accesses and those take the scalar-base + 32-bit
        // vector-offset form; otherwise the extension is hoisted out of the
        // march and every access builds a 64-bit address.  Rewriting the
        // loop-carried copies themselves needs no register copies.
        // LY: thread t stages halo vectors k = t and (LY = 2) k = t + kBlock
        // of the chunk, k < 2H: vector v0 - H + k of the row below the chunk
        // into slot k, or vector v0 + nv + k - H of the row above into slot
        // nv + k; its own vector goes to slot H + t.  Halo vectors outside
        // the plane are zero in both buffers for the whole chunk.
        const long long v0 = (long long)pchunk * kBlock;
        const int nv = (int)min((long long)kBlock, p.nvec_plane - v0);
        const int tid = threadIdx.x, tid2 = tid + kBlock;
        const bool use_h = LY && tid < 2 * H;
        const long long hvec = tid < H ? v0 - H + tid : v0 + nv + (tid - H);
        const bool valid_h = use_h && hvec >= 0 && hvec < p.nvec_plane;
        const int hslot = tid < H ? tid : nv + tid;
        const bool use_h2 = LY == 2 && tid2 < 2 * H; // then tid2 >= H: the row above
        const long long hvec2 = v0 + nv + (tid2 - H);
        const bool valid_h2 = use_h2 && hvec2 < p.nvec_plane;
        const int hslot2 = nv + tid2;
        if constexpr (LY) {
            __syncthreads(); // the previous chunk's last reads are done
            RawV zr;
            __builtin_memset(&zr, 0, sizeof(RawV));
            if (use_h && !valid_h) {
                ybuf[0][hslot] = zr;
                ybuf[1][hslot] = zr;
            }
            if (use_h2 && !valid_h2) {
                ybuf[0][hslot2] = zr;
                ybuf[1][hslot2] = zr;
            }
        }
        // LY: the halo vectors' offsets in a plane take xon's and xos's
        // places (a valid vector when there is nothing to load)
        const uint32_t oh = (uint32_t)((valid_h ? hvec : v0) * V) * (uint32_t)sizeof(T);
        const uint32_t oh2 = (uint32_t)((valid_h2 ? hvec2 : v0) * V) * (uint32_t)sizeof(T);
        uint32_t xoc = oc, xon = LY ? oh : on, xos = LY ? oh2 : os, xob = S::kScaled ? ob >> 1 : ob;
        // stochastic-rounding hash: cell0's part of the pair index times K1
        const uint32_t c0k1 = ((uint32_t)cell0 & ~1u) * kHashK1;
        // FUSED: this wave's row of the slice table and its first cell in a plane
        const uint2 *trow = nullptr;
        uint32_t wcell0 = 0;
        if constexpr (FUSED) {
            const int wv = __builtin_amdgcn_readfirstlane(wave);
            trow = p.wtab + ((long long)pchunk * kWaves + wv) * nz;
            wcell0 = (uint32_t)(((long long)pchunk * kBlock + wv * 64) * V);
        }

        auto issue = [&](int z, Stage &st) {
            const char *Pz = reinterpret_cast<const char *>(P + (long long)z * sxy);
            const char *Pzp = reinterpret_cast<const char *>(P + (long long)z * sxy + (z + 1 < nz ? sxy : 0));
            const char *Qz = reinterpret_cast<const char *>(Q + (long long)z * sxy);
            asm volatile("" : "+v"(xoc), "+v"(xon), "+v"(xos), "+v"(xob));
            const uint32_t xc = xoc;
            st.pp = vload<VT, false>(reinterpret_cast<const VT *>(Pzp + xc));
            st.q = vload<VT, true>(reinterpret_cast<const VT *>(Qz + xc));
            if constexpr (S::kScaled) {
                if constexpr (V == 1)
                    st.cd = load_nibble1(p.code, (long long)z * sxy + cell0);
                else // V divides Nx, so planes and vectors start on even cells
                    st.cd = load_nibbles<V>(p.code + (((long long)z * sxy) >> 1), xob);
            } else {
                st.cd = vload<VB, true>(reinterpret_cast<const VB *>(p.code + (long long)z * sxy + xob));
            }
            if constexpr (LY) {
                st.h = vload<VT, false>(reinterpret_cast<const VT *>(Pz + xon));
                if constexpr (LY == 2) st.h2 = vload<VT, false>(reinterpret_cast<const VT *>(Pz + xos));
            } else {
                st.yn = vload<VT, false>(reinterpret_cast<const VT *>(Pz + xon));
                st.ys = vload<VT, false>(reinterpret_cast<const VT *>(Pz + xos));
                if (need_l) st.tl = reinterpret_cast<const T *>(Pz + xc)[-1];
                if (need_r) st.tr = reinterpret_cast<const T *>(Pz + xc)[V];
            }
            if constexpr (FUSED) {
                // every load goes out every plane, from an index clamped into
                // the slice (one line when the slice is empty), so no load
                // sits behind a branch
                // a scalar load (constant address space): waiting for it
                // counts scalar loads only, not the plane's vector loads
                // issued just before
                using ConstU64 = const __attribute__((address_space(4))) unsigned long long;
                const unsigned long long tv = ((ConstU64 *)(const void *)trow)[z];
                st.efirst = (uint32_t)tv;
                st.ecount = (uint32_t)(tv >> 32);
                const uint32_t k = min((uint32_t)lane, max(st.ecount, 1u) - 1u);
                const long long ec = min((long long)st.efirst + k, p.bs.nb - 1);
                st.eb = p.bidx[ec];
                st.em = p.bs.bmat[ec];
                if constexpr (S::kScaled) {
                    st.es1 = ((const float *)p.bs.ts[0])[ec];
                    st.es2 = ((const float *)p.bs.ts[1])[ec];
                } else {
                    st.ew0 = ((const double *)p.bs.ws[0])[ec];
                    st.ew1 = ((const double *)p.bs.ws[1])[ec];
                }
            }
        };

        C B0[V], B1[V], B2[V];
        Stage SA, SB;
        {
            const VT tm = vload<VT, false>(reinterpret_cast<const VT *>(P + (long long)max(z0 - 1, 0) * sxy + cell0));
            const VT tc = vload<VT, false>(reinterpret_cast<const VT *>(P + (long long)z0 * sxy + cell0));
#pragma unroll
            for (int j = 0; j < V; ++j) {
                B0[j] = z0 >= 1 ? S::ld(tm.v[j]) : (C)0;
                B1[j] = S::ld(tc.v[j]);
            }
            // LY: a body stages P(z) from the other stage's pp, which holds
            // it as stored until the body's issue reloads that stage
            if constexpr (LY) SB.pp = tc;
        }
        issue(z0, SA);

        // One plane: PM/PC hold P(z-1)/P(z); PP receives P(z+1) from cur;
        // nxt receives the loads of plane z+1 before plane z is computed.
        auto body = [&](const int z, C(&PM)[V], C(&PC)[V], C(&PP)[V], Stage &cur, Stage &nxt) {
            C qv[V], yn[V], ys[V];
            uint32_t cd[V];
            // 16-bit: the kw fields of the even cells in the bytes of kwe,
            // of the odd cells in the bytes of kwo, so each converts to float
            // with one byte conversion (the empty asm keeps the compiler
            // from folding the masks back into one extract per cell)
            uint32_t kwe = 0u, kwo = 0u;
            if constexpr (S::kScaled) {
                kwe = cur.cd & 0x07070707u;
                kwo = (cur.cd >> 4) & 0x07070707u;
                asm volatile("" : "+v"(kwe), "+v"(kwo));
            }
            // FP16: the y-neighbours enter the sum straight from their
            // stored pairs (add_f16_half), without a conversion each
            constexpr bool kMixY = std::is_same<T, f16>::value && V >= 2;
            uint32_t ynw[kMixY ? V / 2 : 1], ysw[kMixY ? V / 2 : 1];
            if constexpr (kMixY && !LY) {
                __builtin_memcpy(ynw, &cur.yn, sizeof(VT));
                __builtin_memcpy(ysw, &cur.ys, sizeof(VT));
            }
            // FP16 without a sponge: q enters 2 pc - q straight from its
            // stored pairs as one mixed fused multiply-add by -kq (exact
            // products, one rounding: the bits of the rescale-then-subtract
            // form), so neither a conversion nor a rescale step is needed
            constexpr bool kMixQ = std::is_same<T, f16>::value && V >= 2 && !SPONGE && !FUSED;
            uint32_t qw[kMixQ ? V / 2 : 1];
            if constexpr (kMixQ) __builtin_memcpy(qw, &cur.q, sizeof(VT));
#pragma unroll
            for (int j = 0; j < V; ++j) {
                // at z = nz-1 this is plane z itself; that plane is border,
                // so the value is never used
                PP[j] = S::ld(cur.pp.v[j]);
                if constexpr (!kMixQ) qv[j] = S::ld(cur.q.v[j]);
                if constexpr (!LY) {
                    yn[j] = S::ld(cur.yn.v[j]);
                    ys[j] = S::ld(cur.ys.v[j]);
                }
                if constexpr (S::kScaled)
                    cd[j] = (cur.cd >> (4 * j)) & 0xFu;
                else
                    cd[j] = (uint32_t)cur.cd.v[j];
            }
            if constexpr (S::kScaled && !kMixQ) {
                // Q to R's units; consecutive steps mostly share an exponent,
                // and a multiply by one is exact, so it is skipped then
                if (kq != (C)1)
#pragma unroll
                    for (int j = 0; j < V; ++j) qv[j] = qv[j] * kq;
            }
            C left = __shfl_up(PC[V - 1], 1);
            C right = __shfl_down(PC[0], 1);
            if constexpr (LY) {
                RawV *buf = ybuf[pcount & 1];
                ++pcount;
                RawV r;
                if (active) { // P(z) as stored (nxt.pp, see the prologue)
                    __builtin_memcpy(&r, &nxt.pp, sizeof(VT));
                    buf[H + tid] = r;
                }
                // The last plane of a range issues its own plane again,
                // unused: with a conditional issue the stage values would
                // merge across the branch and be carried through the march,
                // which took 40 more VGPRs (and a wave per SIMD).
                issue(z + 1 < z1 ? z + 1 : z, nxt);
                if (valid_h) {
                    __builtin_memcpy(&r, &cur.h, sizeof(VT));
                    buf[hslot] = r;
                }
                if constexpr (LY == 2)
                    if (valid_h2) {
                        __builtin_memcpy(&r, &cur.h2, sizeof(VT));
                        buf[hslot2] = r;
                    }
                __syncthreads();
                // Lanes past the plane's last vector read slots of other
                // data; their results are not stored, and in 16-bit storage
                // their zero cell code makes each of them +0 or NaN, which
                // the step maximum ignores.
                const RawV rn = buf[2 * H + tid], rs = buf[tid];
                VT ynv, ysv;
                __builtin_memcpy(&ynv, &rn, sizeof(VT));
                __builtin_memcpy(&ysv, &rs, sizeof(VT));
                if constexpr (kMixY) {
                    __builtin_memcpy(ynw, &ynv, sizeof(VT));
                    __builtin_memcpy(ysw, &ysv, sizeof(VT));
                }
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    yn[j] = S::ld(ynv.v[j]);
                    ys[j] = S::ld(ysv.v[j]);
                }
                // the end lanes' x neighbours: cells of the adjacent vectors
                const T *bt = reinterpret_cast<const T *>(buf);
                if (need_l) left = S::ld(bt[(H + tid) * V - 1]);
                if (need_r) right = S::ld(bt[(H + tid + 1) * V]);
            } else {
                if (need_l) left = S::ld(cur.tl);
                if (need_r) right = S::ld(cur.tr);
                if (z + 1 < z1) issue(z + 1, nxt);
            }

            C o[V];
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const uint32_t c = cd[j];
                const C pcv = PC[j];
                if constexpr (S::kScaled) {
                    C sum = (j + 1 < V ? PC[(j + 1) % V] : right) + (j > 0 ? PC[(j + V - 1) % V] : left);
                    if constexpr (kMixY) {
                        if (j & 1) {
                            sum = add_f16_half<true>(sum, ynw[j >> 1]);
                            sum = add_f16_half<true>(sum, ysw[j >> 1]);
                        } else {
                            sum = add_f16_half<false>(sum, ynw[j >> 1]);
                            sum = add_f16_half<false>(sum, ysw[j >> 1]);
                        }
                    } else {
                        sum = sum + yn[j];
                        sum = sum + ys[j];
                    }
                    sum = sum + PP[j];
                    sum = sum + PM[j];
                    const C kw = (C)(((j & 1) ? kwo : kwe) >> (8 * (j >> 1)) & 0xffu); // cells that update
                    // t = sum - kw*pc, u = a*pc - b*q, R = t*(g*c2dt2*ih2) + u.
                    // pc is a 16-bit value (at most 11 significant bits) and kw
                    // an integer up to 6, so kw*pc is exact in FP32 and t as
                    // one fused multiply-add has the bits of the product and
                    // difference rounded apart; likewise 2*pc - q.  With
                    // stochastic rounding R is a fused multiply-add as well.
                    // Nearest rounding keeps R's product rounded apart: its
                    // rounding errors do not average out, and the fused final
                    // step drove the room's constant mode (relative IR error
                    // 0.36 against 5e-3 on the 2000-step shoebox).
                    const C t = __builtin_fmaf(-kw, pcv, sum);
                    C a = two_k, b = (C)1, g = (C)1;
                    if constexpr (SPONGE) {
                        const int d = edge_dist(x0 + j, y, z, nx, ny, nz);
                        if (d < p.sp.w) {
                            a = (C)sp_tab[0][d];
                            b = (C)sp_tab[1][d];
                            g = (C)sp_tab[2][d];
                        }
                    }
                    // Laplacian factor: zero on zero cells when the room has
                    // no hard cells (a*0 - b*0 and t*0 then give +0)
                    const C kf = HARD ? c2ih2 : bsel(bit_mask(c, 3), (C)0, c2ih2);
                    C upd, u2 = (C)0; // u2 = 2 pc - q (no sponge)
                    if constexpr (kMixQ)
                        u2 = (j & 1) ? fma_f16_half<true>(nkq, qw[j >> 1], two_k * pcv)
                                     : fma_f16_half<false>(nkq, qw[j >> 1], two_k * pcv);
                    else if constexpr (!SPONGE)
                        u2 = __builtin_fmaf(two_k, pcv, -qv[j]);
                    if constexpr (SR) {
                        if constexpr (SPONGE)
                            upd = __builtin_fmaf(t, g * kf, __builtin_fmaf(a, pcv, -(b * qv[j])));
                        else
                            upd = __builtin_fmaf(t, kf, u2);
                    } else {
                        if constexpr (SPONGE)
                            upd = (a * pcv - b * qv[j]) + g * (t * kf);
                        else
                            upd = u2 + t * kf;
                    }
                    o[j] = HARD ? bsel(bit_mask(c, 3), bsel(bit_mask(c, 0), hard_u, (C)0), upd) : upd;
                    continue;
                }
                const C xp = bsel(bit_mask(c, 0), pcv, j + 1 < V ? PC[(j + 1) % V] : right);
                const C xm = bsel(bit_mask(c, 1), pcv, j > 0 ? PC[(j + V - 1) % V] : left);
                const C yp = bsel(bit_mask(c, 2), pcv, yn[j]);
                const C ym = bsel(bit_mask(c, 3), pcv, ys[j]);
                const C zp = bsel(bit_mask(c, 4), pcv, PP[j]);
                const C zm = bsel(bit_mask(c, 5), pcv, PM[j]);
                C sum = xp + xm;
                sum = sum + yp;
                sum = sum + ym;
                sum = sum + zp;
                sum = sum + zm;
                const C lap = (sum - (C)6 * pcv) * ih2;
                C a = (C)2, b = (C)1, g = (C)1;
                if constexpr (SPONGE) {
                    const int d = edge_dist(x0 + j, y, z, nx, ny, nz);
                    if (d < p.sp.w) {
                        a = (C)sp_tab[0][d];
                        b = (C)sp_tab[1][d];
                        g = (C)sp_tab[2][d];
                    }
                }
                const C upd = (a * pcv - b * qv[j]) + g * (c2dt2 * lap);
                // type bits 6-7: 00 update, 11 update + filter, 10 hard source, 01 zero
                const uint32_t differ = bit_mask(c ^ (c >> 1), 6);
                const uint32_t is_hard = differ & bit_mask(c, 7);
                o[j] = bsel(differ, bsel(is_hard, hard_u, (C)0), upd);
            }

            VT out;
            if constexpr (S::kScaled) {
#pragma unroll
                for (int j = 0; j < V; ++j) lmax = fmaxf(lmax, fabsf((float)o[j]));
                if constexpr (SR) {
                    // the hash of mix16 for cell i = z*sxy + cell0 in 32-bit
                    // pieces: hk takes the high word of i (the plane's plus a
                    // carry), lo the pair's low word times K1
                    const unsigned long long zs = (unsigned long long)z * (unsigned long long)sxy;
                    const uint32_t zlo = (uint32_t)zs, ilo = zlo + (uint32_t)cell0;
                    const uint32_t hk0 = hash_hk(zs, n, kSaltDense);
                    const uint32_t hk = ilo < zlo ? hk0 + kHashK3 : hk0;
                    if constexpr (V == 1) {
                        const uint32_t h = hash_pair((ilo & ~1u) * kHashK1, hk);
                        const float q0 = sr_bits<T>(o[0], (ilo & 1u) ? (h >> 16) : (h & 0xffffu));
                        T dummy;
                        cvt2_truncated<T>(q0, 0.f, out.v[0], dummy);
                    } else {
                        // z*sxy and cell0 are even here
                        const uint32_t lo = zlo * kHashK1 + c0k1;
#pragma unroll
                        for (int j = 0; j < V; j += 2) {
                            // both halves masked at once (the same bits as
                            // masking each 16-bit half); the empty asm keeps
                            // the compiler from splitting the mask per half
                            uint32_t h = hash_pair(lo + (uint32_t)j * kHashK1, hk) & (sr_mask<T>() * 0x10001u);
                            asm volatile("" : "+v"(h));
                            const float qa = __uint_as_float(__float_as_uint(o[j]) + (h & 0xffffu));
                            const float qb = __uint_as_float(__float_as_uint(o[j + 1]) + (h >> 16));
                            cvt2_truncated<T>(qa, qb, out.v[j], out.v[j + 1]);
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < V; ++j) out.v[j] = S::st(o[j]);
                }
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j) out.v[j] = S::st(o[j]);
            }
            if constexpr (FUSED) {
                // Boundary phase (simulator_gpu.cu:126-148): the list entries
                // of this wave's cells at plane z.  Each cell publishes its
                // filter input d = p^n - p^{n-1} (physical units) and its
                // stored p^{n+1}; lane L takes entry first + L (64 per round),
                // reads its cell's pair, runs the DF1 step and the correction
                // exactly as the list launch does, stores the filter state
                // (coalesced: consecutive entries on consecutive lanes) and
                // writes the corrected stored value back for the cell's lane
                // to store with the rest of its vector.
                const uint32_t cnt = cur.ecount;
                if (cnt != 0u) {
#pragma unroll
                    for (int j = 0; j < V; ++j) {
                        const C d = S::kScaled ? PC[j] * kphys - qv[j] * kst_inv : PC[j] - qv[j];
                        bslot[wave][j][lane] = BSlot{d, S::ld(out.v[j])};
                    }
                    wave_lds_fence();
                    const uint32_t pbase = (uint32_t)((unsigned long long)z * (unsigned long long)sxy) + wcell0;
                    const bool odd = (n & 1) != 0;
                    for (uint32_t r0 = 0; r0 < cnt; r0 += 64) {
                        uint32_t b = cur.eb, m = cur.em;
                        double wa = cur.ew0, wb = cur.ew1;
                        float sa = cur.es1, sb = cur.es2;
                        const long long e = (long long)cur.efirst + r0 + lane;
                        if (r0 > 0) { // entries past the first 64 (floor and ceiling planes)
                            const long long ec = min(e, (long long)cur.efirst + cnt - 1);
                            b = p.bidx[ec];
                            m = p.bs.bmat[ec];
                            if constexpr (S::kScaled) {
                                sa = ((const float *)p.bs.ts[0])[ec];
                                sb = ((const float *)p.bs.ts[1])[ec];
                            } else {
                                wa = ((const double *)p.bs.ws[0])[ec];
                                wb = ((const double *)p.bs.ws[1])[ec];
                            }
                        }
                        if ((uint32_t)lane < cnt - r0) {
                            const uint32_t off = b - pbase;
                            BSlot &sl = bslot[wave][off % (uint32_t)V][off / (uint32_t)V];
                            const C dif = sl.d;
                            const double in = (double)dif;
                            const double *cf = cf_lds + 5 * m;
                            double yv, wv = 0.0, s1 = 0.0, s2 = 0.0;
                            if constexpr (S::kScaled) {
                                yv = tdf2_out(cf, in, (double)sa);
                                s1 = tdf2_s1(cf, in, yv, (double)sb);
                                s2 = tdf2_s2(cf, in, yv);
                            } else {
                                const double w1 = odd ? wa : wb, w2 = odd ? wb : wa;
                                wv = df2_state(cf, in, w1, w2);
                                yv = df2_out(cf, wv, w1, w2);
                            }
                            const C r = corrected<C>(sl.r * kst_inv, p.bs.kconst, yv, in);
                            flag_nonfinite(p.ctl, r);
                            const T t = store_round<T>(r * kst, SR, (unsigned long long)b, n, kSaltBoundary);
                            if constexpr (S::kScaled) store_tdf2(p.bs, e, s1, s2);
                            else store_df2<C>(p.bs, e, n, dif, yv, wv);
                            sl.r = S::ld(t);
                        }
                    }
                    wave_lds_fence();
#pragma unroll
                    for (int j = 0; j < V; ++j) out.v[j] = S::st(bslot[wave][j][lane].r);
                }
            }
            if (active) vstore<VT, true>(reinterpret_cast<VT *>(reinterpret_cast<char *>(R + (long long)z * sxy) + xoc), out);
        };

        // planes rotate with period 3, load stages with period 2
        for (int z = z0; z < z1; z += 6) {
            body(z, B0, B1, B2, SA, SB);
            if (z + 1 >= z1) break;
            body(z + 1, B1, B2, B0, SB, SA);
            if (z + 2 >= z1) break;
            body(z + 2, B2, B0, B1, SA, SB);
            if (z + 3 >= z1) break;
            body(z + 3, B0, B1, B2, SB, SA);
            if (z + 4 >= z1) break;
            body(z + 4, B1, B2, B0, SA, SB);
            if (z + 5 >= z1) break;
            body(z + 5, B2, B0, B1, SB, SA);
        }
    }

    if constexpr (S::kScaled) {
        __shared__ float red[kWaves];
        lmax = wave_max(lmax) * (float)kst_inv; // R's units -> physical
        if (lane == 0) red[wave] = lmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = red[0];
            for (int k = 1; k < kWaves; ++k) m = fmaxf(m, red[k]);
            atomicMax(&p.ctl->maxabs[n & 3][bid % kSlots], __float_as_uint(m));
        }
    }
}

struct TailParams {
    const void *P;
    const void *Q;
    void *R;
    const uint32_t *bidx;    // compact list; bs.nb = 0 when the stencil fused it
    BoundaryState bs;
    // Stored pressures of the list cells as a two-slot ring by step parity:
    // before step n, hv[n&1] holds p^n and hv[(n+1)&1] p^{n-1} at each
    // entry's cell, in the field's storage type; step n replaces p^{n-1} by
    // the p^{n+1} it stores.  Null: read p^n and p^{n-1} from the fields.
    void *hv[2];
    StepCtl *ctl;
    int slot, next_slot;     // reads nring[slot], writes nring[next_slot]
    int nsoft;
    const long long *soft_idx, *soft_bpos;
    const double *soft_sig;
    long long soft_len;
    int nrcv;
    const long long *rcv_idx, *rcv_bpos;
    // per receiver k: the soft source on its cell (-1: none), the next
    // receiver on the same cell (-1: none), 1 if no earlier receiver shares
    // it; per soft source: the first receiver on its cell (-1: none)
    const int *rcv_src, *rcv_next, *rcv_first, *src_rcv;
    int list_points;         // some soft source or receiver sits on a list cell
    double *ir;              // forward: [nrcv][ir_cap] output
    const double *gsig;      // reverse: dL/dIR [nrcv][ir_cap]
    long long ir_cap;
    double *grad_part;       // reverse: [gridDim][nmat][5]
    int nx, ny, nz;
    Sponge sp;
    int sr;                  // stochastic storage rounding (16-bit formats)
};

template <typename C>
__device__ inline void sponge_g(const TailParams &p, long long idx, C &g)
{
    g = (C)1;
    if (p.sp.w > 0) {
        const long long sxy = (long long)p.nx * p.ny;
        const int z = (int)(idx / sxy);
        const long long r = idx - (long long)z * sxy;
        const int y = (int)(r / p.nx);
        const int x = (int)(r - (long long)y * p.nx);
        const int d = edge_dist(x, y, z, p.nx, p.ny, p.nz);
        if (d < p.sp.w) g = (C)p.sp.g[d];
    }
}

// Distance of a list cell from the domain edge (the sponge table index):
// list indices are 32-bit, so the coordinates come from 32-bit divisions
// (exact, and far shorter than 64-bit ones).
__device__ inline int sponge_dist32(const TailParams &p, uint32_t idx)
{
    const uint32_t nx = (uint32_t)p.nx, sxy = nx * (uint32_t)p.ny;
    const uint32_t z = idx / sxy;
    const uint32_t r = idx - z * sxy;
    const uint32_t y = r / nx;
    const uint32_t x = r - y * nx;
    return edge_dist((int)x, (int)y, (int)z, p.nx, p.ny, p.nz);
}

// List entries per thread of the forward list launch.  One is fastest for
// every precision on the 512x384x256 room (FP16 12.3 / 13.5 / 14.1 us and
// FP32 15.8 / 16.8 / 18.3 us at 1 / 2 / 4 entries).
#ifndef VIRS_TAIL_ENTRIES
#define VIRS_TAIL_ENTRIES 1
#endif
constexpr int kTailEntries = VIRS_TAIL_ENTRIES;
// Threads per block of the forward list launch.  On the 512x384x256 room
// 128 is fastest: FP16 714 against 708-710 Gvox/s per step with 256 (list
// launch 12.0-12.4 against 12.3-12.9 us), FP32 404 against 401; 512 and
// 1024 are slower, 64 the same or slightly slower (FP16 714-715 against
// 717-718).  Block 0's point work strides over its threads, so any whole
// number of waves works.
#ifndef VIRS_TAIL_BLOCK
#define VIRS_TAIL_BLOCK 128
#endif
constexpr int kTailBlock = VIRS_TAIL_BLOCK;
static_assert(kTailBlock >= 64 && kTailBlock % 64 == 0, "whole waves");
// Threads per block of the reverse list launch (one row of gradient
// partials per wave).  128 and 256 measure alike (FP32 10.55 / 10.61 us on
// the 256x192x128 room).
#ifndef VIRS_TAIL_ADJ_BLOCK
#define VIRS_TAIL_ADJ_BLOCK 256
#endif
constexpr int kTailAdjBlock = VIRS_TAIL_ADJ_BLOCK;
constexpr int kTailAdjWaves = kTailAdjBlock / 64;
static_assert(kTailAdjBlock >= 64 && kTailAdjBlock % 64 == 0, "whole waves");

// The rest of a step after k_dense: the boundary biquad over the compact
// list when it was not fused into the stencil, then soft sources (added to
// the corrected field), receiver gather, and the step counter.  Point cells
// on the list are handled by their entry's thread, the others by block 0.
//
// One entry per thread (kTailEntries).  The launch is short and its p^{n+1}
// accesses are scattered, so latency weighs as much as bytes: every load
// that the entry index e names
// (the list index, the material byte, both slots of the pressure ring hv and
// of the filter taps) is issued at once, from e clamped into the list, and
// none depends on the step counter: both ring slots are read every step, and
// the counter only picks which slot is which.  What is left in series is
// the list index, then the gather of p^{n+1} at the entry's cell; the
// material's coefficients (cached) follow the material byte in parallel.
// No load sits behind a branch, since a guarded load makes the compiler wait
// for it where the branch rejoins.  The stores are predicated on the entry
// being in the list and the step not skipped (stop bit).
template <typename T>
__global__ __launch_bounds__(kTailBlock) void k_tail(const TailParams p)
{
    using S = Store<T>;
    using C = typename S::C;
    // The step (with the stop bit) and, for 16-bit fields, the storage
    // scales of p^n, p^{n-1} and p^{n+1}.  The list blocks read them after
    // issuing their entries' loads, which do not depend on them, so the
    // counter's load overlaps those instead of preceding them.
    struct StepState {
        long long n;
        bool skip, in_sig, in_ir;
        C sP, sQ, sR, sRi;
    };
    auto step_state = [&]() {
        StepState st;
        st.n = p.ctl->nring[p.slot];
        st.skip = (st.n & kStopBit) != 0;
        st.in_sig = st.n >= 0 && st.n < p.soft_len;
        st.in_ir = st.n >= 0 && st.n < p.ir_cap;
        st.sP = st.sQ = st.sR = st.sRi = (C)1;
        if constexpr (S::kScaled) {
            // all four exponents in one load, selected by the step
            const int4 ex = *(const int4 *)p.ctl->exp;
            const int ring[4] = {ex.x, ex.y, ex.z, ex.w};
            st.sP = exp2i(-ring[(st.n - 1) & 3]);
            st.sQ = exp2i(-ring[(st.n - 2) & 3]);
            const int eR = ring[st.n & 3];
            st.sR = exp2i(eR);
            st.sRi = exp2i(-eR);
        }
        return st;
    };
    T *__restrict__ R = (T *)p.R;
    const long long nb = p.bs.nb;
    if (blockIdx.x > 0) { // list blocks exist only when nb > 0
        constexpr int U = kTailEntries;
        // block b covers entries [(b-1)*U*kTailBlock, b*U*kTailBlock), lane-contiguous
        // per slot u; all loads of all U entries go out before the gathers
        long long ev[U];
        uint32_t bv[U];
        uint8_t mv[U];
        T h0[U], h1[U], rt[U];
        // filter state: direct form II (FP64 w ring) for FP32/FP64 fields,
        // transposed direct form II (FP32 s1, s2) for 16-bit fields
        constexpr bool kDF2 = !S::kScaled;
        double wa[U], wb[U];
        float sa[U], sb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ev[u] = ((long long)(blockIdx.x - 1) * U + u) * kTailBlock + threadIdx.x;
            const long long ec = ev[u] < nb ? ev[u] : nb - 1;
            bv[u] = p.bidx[ec];
            mv[u] = p.bs.bmat[ec]; // always present for the list launch
            h0[u] = ((const T *)p.hv[0])[ec];
            h1[u] = ((const T *)p.hv[1])[ec];
            if constexpr (kDF2) {
                wa[u] = ((const double *)p.bs.ws[0])[ec];
                wb[u] = ((const double *)p.bs.ws[1])[ec];
            } else {
                sa[u] = ((const float *)p.bs.ts[0])[ec];
                sb[u] = ((const float *)p.bs.ts[1])[ec];
            }
        }
        const StepState ss = step_state();
        const long long n = ss.n;
        const bool skip = ss.skip, in_sig = ss.in_sig, in_ir = ss.in_ir;
        const C sP = ss.sP, sQ = ss.sQ, sR = ss.sR, sRi = ss.sRi;
#pragma unroll
        for (int u = 0; u < U; ++u) rt[u] = R[bv[u]];
        // before step n, tap slot n&1 holds step n-2's value and slot
        // (n+1)&1 step n-1's (BoundaryState); hv[n&1] holds p^n and
        // hv[(n+1)&1] p^{n-1} (TailParams::hv)
        const bool odd = (n & 1) != 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long e = ev[u];
            const uint32_t b = bv[u];
            const double *cf = p.bs.coeffs + 5 * mv[u];
            C pp = S::ld(odd ? h1[u] : h0[u]), qq = S::ld(odd ? h0[u] : h1[u]), r = S::ld(rt[u]);
            if constexpr (S::kScaled) pp = pp * sP, qq = qq * sQ, r = r * sRi;
            const C dif = pp - qq;
            const double in = (double)dif;
            double yv, wv = 0.0, s1 = 0.0, s2 = 0.0;
            if constexpr (kDF2) {
                const double w1 = odd ? wa[u] : wb[u], w2 = odd ? wb[u] : wa[u];
                wv = df2_state(cf, in, w1, w2);
                yv = df2_out(cf, wv, w1, w2);
            } else {
                yv = tdf2_out(cf, in, (double)sa[u]);
                s1 = tdf2_s1(cf, in, yv, (double)sb[u]);
                s2 = tdf2_s2(cf, in, yv);
            }
            r = corrected<C>(r, p.bs.kconst, yv, in);
            T t = store_round<T>(r * sR, p.sr, (unsigned long long)b, n, kSaltBoundary);
            if (e < nb && !skip) {
                flag_nonfinite(p.ctl, r);
                if constexpr (kDF2) store_df2<C>(p.bs, e, n, dif, yv, wv);
                else store_tdf2(p.bs, e, s1, s2);
                // a soft source on this cell is added to the stored corrected
                // value (the point tables are scanned only when some point is
                // on the list)
                for (int k = 0; k < (p.list_points ? p.nsoft : 0); ++k)
                    if (p.soft_bpos[k] == e && in_sig) {
                        C r2 = S::ld(t);
                        if constexpr (S::kScaled) r2 = r2 * sRi;
                        r2 = r2 + (C)p.soft_sig[(long long)k * p.soft_len + n];
                        t = store_round<T>(r2 * sR, p.sr, (unsigned long long)b, n, kSaltSource);
                    }
                R[b] = t;
                ((T *)p.hv[(n + 1) & 1])[e] = t;
                if (in_ir && p.list_points)
                    for (int k = 0; k < p.nrcv; ++k)
                        if (p.rcv_bpos[k] == e) {
                            C stored = S::ld(t);
                            if constexpr (S::kScaled) stored = stored * sRi;
                            flag_nonfinite(p.ctl, stored);
                            p.ir[(long long)k * p.ir_cap + n] = (double)stored;
                        }
            }
        }
        return;
    }
    const StepState ss = step_state();
    const long long n = ss.n;
    const bool in_sig = ss.in_sig, in_ir = ss.in_ir;
    const C sR = ss.sR, sRi = ss.sRi;
    // Block 0: the points off the list, the RMS rule and the step counter.
    // Every table load is issued at once.  While a soft source is active,
    // its thread also writes the samples of the receivers on its cell, so
    // no receiver reads a cell that another thread of the launch writes
    // (and no barrier orders the two).
    const int k = threadIdx.x;
    StepCtl *ctl = p.ctl;
    const bool rms = ctl->es_enable && p.nrcv > 0 && n > kRmsMinStep && n <= p.ir_cap;
    __shared__ float win[kRmsWindow];
    for (int q = k; rms && q < kRmsWindow; q += kTailBlock) win[q] = (float)p.ir[n - kRmsWindow + q];
    for (int ks = k; ks < p.nsoft && in_sig; ks += kTailBlock) {
        const long long i = p.soft_idx[ks], bp = p.soft_bpos[ks];
        const int j0 = p.src_rcv[ks];
        const C sg = (C)p.soft_sig[(long long)ks * p.soft_len + n];
        C r = S::ld(R[i]);
        if constexpr (S::kScaled) r = r * sRi;
        r = r + sg;
        const T t = store_round<T>(r * sR, p.sr, (unsigned long long)i, n, kSaltSource);
        if (bp < 0) {
            R[i] = t;
            if (in_ir) {
                C v = S::ld(t);
                if constexpr (S::kScaled) v = v * sRi;
                if (j0 >= 0) flag_nonfinite(ctl, v);
                for (int j = j0; j >= 0; j = p.rcv_next[j]) p.ir[(long long)j * p.ir_cap + n] = (double)v;
            }
        }
    }
    for (int kr = k; kr < p.nrcv && in_ir; kr += kTailBlock) {
        const long long i = p.rcv_idx[kr], bp = p.rcv_bpos[kr];
        const bool by_source = in_sig && p.rcv_src[kr] >= 0;
        C r = S::ld(R[i]);
        if constexpr (S::kScaled) r = r * sRi;
        if (bp < 0 && !by_source) {
            flag_nonfinite(ctl, r);
            p.ir[(long long)kr * p.ir_cap + n] = (double)r;
        }
    }
    // RMS rule of step n over receiver 0's samples [n-100, n), in float in
    // the reference's order (simulator_gpu.cu:303-330)
    long long next = n + 1;
    if (rms) {
        __syncthreads();
        if (k == 0) {
            float sum = 0.0f;
            for (int i = 0; i < kRmsWindow; ++i) sum += win[i] * win[i];
            float rms_v = __fsqrt_rn(sum / (float)kRmsWindow);
            if (n == kRmsMinStep + 1) {
                ctl->rms_state = rms_v;
            } else {
                const float fg = 0.00008f;
                const float vout = (fg * rms_v + ctl->rms_state) / (fg + 1.f);
                ctl->rms_state = 2.f * vout - ctl->rms_state;
                rms_v = vout;
            }
            ctl->percent = ctl->es_target / rms_v;
            if (rms_v < ctl->es_target) {
                ctl->stop_at = n;
                next |= kStopBit;
            }
        }
    }
    if (k == 0) ctl->nring[p.next_slot] = next;
}

// Reverse boundary step.  With w = g*rbar the adjoint of the field after
// k_dense<ADJ> (g = sponge factor of the cell), for each entry:
//   rbar   = w/g + dbar1 - dbar2 (+ dL/dIR if a receiver sits here)
//   ybar   = -0.1 kap rbar - a1 ybar1 - a2 ybar2          kap = K*0.4f
//   dbar   = -0.9 kap rbar + b0 ybar + b1 ybar1 + b2 ybar2
//   dL/d(b0,b1,b2,a1,a2) += ybar * (d^n, d^{n-1}, d^{n-2}, -y^{n-1}, -y^{n-2})
// where d and y come from the forward trace of this segment.  The state
// rings of BoundaryState hold ybar (ys) and dbar (xs).  Per-wave gradient
// partials are summed in a fixed order, so the result is deterministic.  (A
// per-block sum through LDS put a barrier, which waits for the waves'
// stores, at the end of every block: 17.0 against 10.0 us per launch.)
template <typename T>
__global__ __launch_bounds__(kTailAdjBlock) void k_tail_adj(const TailParams p)
{
    using C = T;
    const long long n = p.ctl->nring[p.slot];
    const bool in_g = n >= 0 && n < p.ir_cap;
    T *__restrict__ R = (T *)p.R;
    const BoundaryState &bs = p.bs;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (blockIdx.x == 0) {
        // Block 0: dL/dIR into the adjoint field at the receivers off the
        // list, and the step counter.  Receivers that share a cell are
        // summed by the first of them, in index order, so no two threads
        // update one cell.
        const int k = threadIdx.x;
        if (k < p.nrcv && in_g && p.rcv_bpos[k] < 0 && p.rcv_first[k]) {
            const long long i = p.rcv_idx[k];
            double g = 1.0;
            sponge_g(p, i, g);
            C acc = (C)((double)R[i] + g * p.gsig[(long long)k * p.ir_cap + n]);
            for (int j = p.rcv_next[k]; j >= 0; j = p.rcv_next[j])
                acc = (C)((double)acc + g * p.gsig[(long long)j * p.ir_cap + n]);
            R[i] = acc;
        }
        if (k == 0) p.ctl->nring[p.next_slot] = n - 1;
        return;
    }
    // list blocks (they exist only when the list is not empty)
    const long long e = (long long)(blockIdx.x - 1) * kTailAdjBlock + threadIdx.x;
    // this wave's running sums (one row per wave, so no barrier joins the
    // waves), loaded now so the load is not the last link of the chain.
    // With at most 16 sums (tree reduction below) lane 4s writes sum s,
    // otherwise lane s.
    const bool tree = 5 * bs.nmat <= 16;
    const int pslot = tree ? ((lane & 3) == 0 ? lane >> 2 : 64) : lane;
    const bool writer = pslot < 5 * bs.nmat;
    double *dst = p.grad_part + ((long long)(blockIdx.x - 1) * kTailAdjWaves + wave) * 5 * bs.nmat + (writer ? pslot : 0);
    const double prev = writer ? *dst : 0.0;
    double gv[5] = {0, 0, 0, 0, 0};
    int m = -1;
    const long long nb = bs.nb;
    {
        // as in k_tail: every load that e names is issued at once from e
        // clamped into the list, and the stores are predicated
        const long long ec = e < nb ? e : nb - 1;
        const uint32_t b = p.bidx[ec];
        const int mm = bs.bmat[ec];
        const int sp = (int)(n & 1);
        const double xa = ((const double *)bs.xs[0])[ec], xb = ((const double *)bs.xs[1])[ec];
        const double ya = ((const double *)bs.ys[0])[ec], yc = ((const double *)bs.ys[1])[ec];
        const long long sl = n - bs.trace_base + 2;
        const T *td = (const T *)bs.trace_d, *ty = (const T *)bs.trace_y;
        const T tdn = td[sl * nb + ec], tx1 = td[(sl - 1) * nb + ec], tx2 = td[(sl - 2) * nb + ec];
        const T ty1 = ty[(sl - 1) * nb + ec], ty2 = ty[(sl - 2) * nb + ec];
        const T rt = R[b];
        // sponge factor without a branch: the table entry is read for every
        // cell (clamped into the table) and used where the cell lies in the
        // sponge; x/1 and x*1 are exact, so rooms without a sponge get the
        // same bits as skipping both
        const int sd = sponge_dist32(p, b);
        const float gt = p.sp.g[sd < kMaxSponge ? sd : kMaxSponge - 1];
        const double *cf = bs.coeffs + 5 * mm;
        const double c0 = cf[0], c1 = cf[1], c2 = cf[2], c3 = cf[3], c4 = cf[4];
        // every load above is issued before any arithmetic on what they
        // return (left to itself the compiler waits for the gather first)
        __builtin_amdgcn_sched_barrier(0);
        const double dn = (double)tdn, x1 = (double)tx1, x2 = (double)tx2;
        const double y1 = (double)ty1, y2 = (double)ty2;
        const double g = sd < p.sp.w ? (double)gt : 1.0;
        double rb = (double)rt / g;
        // slot sp holds step n+2's adjoint taps, slot sp^1 step n+1's
        const double yb1 = sp ? ya : yc, yb2 = sp ? yc : ya;
        const double db1 = sp ? xa : xb, db2 = sp ? xb : xa;
        rb = rb + db1 - db2;
        if (in_g && p.list_points)
            for (int k = 0; k < p.nrcv; ++k)
                if (p.rcv_bpos[k] == e) rb += p.gsig[(long long)k * p.ir_cap + n];
        const double kap = (double)bs.kconst * (double)0.4f;
        const double yb = -0.1 * kap * rb - c3 * yb1 - c4 * yb2;
        const double db = -0.9 * kap * rb + c0 * yb + c1 * yb1 + c2 * yb2;
        if (e < nb) {
            m = mm;
            gv[0] = yb * dn;
            gv[1] = yb * x1;
            gv[2] = yb * x2;
            gv[3] = -yb * y1;
            gv[4] = -yb * y2;
            ((double *)bs.ys[sp])[e] = yb;
            ((double *)bs.xs[sp])[e] = db;
            R[b] = (C)(rb * g);
        }
    }
    double acc = 0.0;
    if (tree) {
        // Up to three materials: the 5*nmat sums of the wave in one
        // transposed butterfly over 16 slots (slot 5*m + c holds this lane's
        // term of coefficient c of material m).  At each of the first four
        // steps a lane keeps half of its slots, adds its partner's copies of
        // them and hands over the other half, so 8+4+2+1 exchanges leave lane
        // L with the partial of slot (L >> 2) & 15; two plain steps finish it.
        // Fixed order, so the sums are deterministic.
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = 0.0;
#pragma unroll
        for (int c = 0; c < 5; ++c)
#pragma unroll
            for (int mm = 0; mm < 3; ++mm)
                if (m == mm) v[5 * mm + c] = gv[c];
#pragma unroll
        for (int h = 8, d = 32; h >= 1; h >>= 1, d >>= 1) {
            const bool up = (lane & d) != 0;
#pragma unroll
            for (int q = 0; q < h; ++q) {
                const double keep = up ? v[h + q] : v[q];
                const double give = up ? v[q] : v[h + q];
                v[q] = keep + __shfl_xor(give, d);
            }
        }
        double t = v[0];
        t = t + __shfl_xor(t, 2);
        t = t + __shfl_xor(t, 1);
        if (writer) *dst = prev + t; // lane 4s holds the sum of slot s
        return;
    }
    // More materials: per-material wave sums, fixed reduction order.  Lanes
    // without an entry hold m = -1 and zero terms; a wave whose entries share
    // one material reduces only that material's five sums (the others are
    // the sums of zeros either way, so the result is the same bit for bit).
    // Every lane ends with every sum; lane 5*mm + c adds sum (mm, c) to the
    // wave's row.  The host adds the rows in a fixed order.
    const int mw = wave_max_int(m);
    const bool uniform = __ballot(m >= 0 && m != mw) == 0ull;
    if (uniform) {
        if (mw >= 0) {
            double sum[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) sum[c] = wave_sum(gv[c]);
#pragma unroll
            for (int c = 0; c < 5; ++c)
                if (lane == mw * 5 + c) acc = sum[c];
        }
    } else {
        for (int mm = 0; mm < bs.nmat; ++mm) {
            double sum[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) sum[c] = wave_sum(m == mm ? gv[c] : 0.0);
#pragma unroll
            for (int c = 0; c < 5; ++c)
                if (lane == mm * 5 + c) acc = sum[c];
        }
    }
    if (writer) *dst = prev + acc;
}

// Seeds the ring of list-cell pressures from the fields before step n:
// pn = hv[n&1] receives p^n, pq = hv[(n+1)&1] receives p^{n-1}.
template <typename T>
__global__ void k_stash(const T *__restrict__ P, const T *__restrict__ Q, const uint32_t *__restrict__ bidx,
                        T *__restrict__ pn, T *__restrict__ pq, long long nb)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nb;
         e += (long long)gridDim.x * blockDim.x) {
        const uint32_t b = bidx[e];
        pn[e] = P[b];
        pq[e] = Q[b];
    }
}

// Converts a stored field to double in physical units (value * 2^-e).
template <typename T>
__global__ void k_to_f64(const T *__restrict__ in, double *__restrict__ out, long long n, int e)
{
    using S = Store<T>;
    const double s = (double)exp2i(-e);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (double)S::ld(in[i]) * s;
}

template <typename T>
__global__ void k_from_f64(const double *__restrict__ in, T *__restrict__ out, long long n, int e)
{
    using S = Store<T>;
    using C = typename S::C;
    const C s = (C)exp2i(e);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = S::st((C)in[i] * s);
}

// Packs the XZ plane y = layer of a stored field into out[z*nx + x] as float
// in physical units, converting exactly as the whole-field readback does
// (through double, value * 2^-e).  For scaled storage the exponent is read on
// the device from the step counter of rotation phase `slot`, so the launch
// can queue behind steps whose exponents the host has not seen.
template <typename T>
__global__ void k_plane_y(const T *__restrict__ f, const StepCtl *__restrict__ ctl, int slot,
                          float *__restrict__ out, int nx, int ny, int nz, int layer)
{
    using S = Store<T>;
    int e = 0;
    if constexpr (S::kScaled) e = ctl->exp[((ctl->nring[slot] & ~kStopBit) - 1) & 3];
    const double s = (double)exp2i(-e);
    for (int z = blockIdx.y; z < nz; z += gridDim.y) {
        const long long row = ((long long)z * ny + layer) * nx;
        for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < nx; x += gridDim.x * blockDim.x)
            out[(long long)z * nx + x] = (float)((double)S::ld(f[row + x]) * s);
    }
}

// Sets the step index of rotation phase `slot` (and, with reset, clears the
// control block's step state; the early-stop settings stay) without a
// memset, so it can sit inside a captured graph.
__global__ void k_ctl_set(StepCtl *ctl, int slot, long long n, int reset)
{
    const int t = threadIdx.x;
    if (reset) {
        if (t < 4) {
            ctl->exp[t] = 0;
            for (int s = 0; s < kSlots; ++s) ctl->maxabs[t][s] = 0u;
        }
        if (t == 0) {
            ctl->err = 0;
            ctl->stop_at = 0x7fffffffffffffffll;
            ctl->rms_state = 0.f;
            ctl->percent = 0.f;
        }
    }
    if (t == 0) ctl->nring[slot] = n;
}

// Early-stop settings of the control block.
__global__ void k_ctl_es(StepCtl *ctl, int enable, float target)
{
    if (threadIdx.x == 0) {
        ctl->es_enable = enable;
        ctl->es_target = target;
    }
}

// Copies the filter state at a segment start into trace slots 0 (step s0-2)
// and 1 (step s0-1): x2,y2 and x1,y1.
template <typename C>
__global__ void k_trace_seed(const C *x1, const C *x2, const double *y1, const double *y2,
                             C *trace_d, C *trace_y, long long nb)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nb;
         e += (long long)gridDim.x * blockDim.x) {
        trace_d[e] = x2[e];
        trace_y[e] = (C)y2[e];
        trace_d[nb + e] = x1[e];
        trace_y[nb + e] = (C)y1[e];
    }
}

} // namespace virs
