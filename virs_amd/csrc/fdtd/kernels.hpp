// Device kernels of the FDTD step for gfx950 (CDNA4, wave64).
//
// Semantics follow the reference GPU step (src/simulator/simulator_gpu.cu):
//   dense update      :38-103   -> k_dense
//   boundary biquad   :105-150  -> k_tail (with point sources, receivers and
//                                  the device step counter)
//   adjoint           (new)     -> k_dense<ADJ> + k_tail_adj
// Arithmetic is written in the reference's operand order and the file is
// compiled without FP contraction, so an FP32/FP64 run is bit-identical to a
// CPU evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace virs {

constexpr int kBlock = 256;   // 4 waves of 64
constexpr int kSlots = 64;    // spread of the per-step |p| maximum atomics
constexpr int kMaxSponge = 64;
constexpr int kMaxPoints = 64;

// Device-resident step state.  n is the step being computed; exp/maxabs are
// rings indexed by step & 3 holding each stored field's power-of-two scale
// exponent and its max |value| (physical units, float bits; non-negative
// floats order like their bit patterns).
struct StepCtl {
    long long n;
    unsigned int done;
    int err;
    int exp[4];
    unsigned int maxabs[4][kSlots];
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

template <typename T, int V>
struct alignas(sizeof(T) * V) Vec {
    T v[V];
};

__device__ inline float exp2i(int e) { return __int_as_float((e + 127) << 23); }

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

// Cell code, one byte per cell, built on the device from the reference flags:
//   bits 0-5  neighbour is a wall (flag 1): +x 1, -x 2, +y 4, -y 8, +z 16, -z 32
//   bits 6-7  0 update (flags 0 and 2), 1 write zero (flag 1 and the domain
//             border), 2 hard source (flag 3)
// The stencil then reads one byte per cell instead of seven flags.
enum : uint8_t { kTypeUpdate = 0, kTypeZero = 1, kTypeHard = 2 };

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
                c = 0;
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

struct Sponge {
    int w;
    float a[kMaxSponge], b[kMaxSponge], g[kMaxSponge];
};

__device__ inline int edge_dist(int x, int y, int z, int nx, int ny, int nz)
{
    return min(min(min(x, nx - 1 - x), min(y, ny - 1 - y)), min(z, nz - 1 - z));
}

struct DenseParams {
    const void *P;   // p^n      (forward) / adjoint of p^{n+1} (reverse)
    const void *Q;   // p^{n-1}  (forward) / adjoint of p^{n+2} (reverse)
    void *R;         // p^{n+1}  (forward) / adjoint of p^n     (reverse)
    const uint8_t *code;
    int nx, ny, nz;
    long long sxy;
    long long nvec_plane; // vectors of V cells per z-plane
    int nchunks_plane;    // blocks per z-plane
    int zc;               // z-planes per block
    float c2dt2, ih2;
    StepCtl *ctl;
    const double *hard_sig;
    long long hard_len;
    int hard_ref;         // 1: the reference pulse (10 for n < 10)
    int nsoft;            // soft sources: their |signal| enters the scale estimate
    const double *soft_sig;
    long long soft_len;
    Sponge sp;
};

// Leapfrog 7-point update, simulator_gpu.cu:38-103 with a graded sponge:
//   lap = (V(x+1)+V(x-1)+V(y+1)+V(y-1)+V(z+1)+V(z-1) - 6 pc) * ih2
//   R   = (a pc - b Q) + g (c2dt2 lap),   V(nbr) = pc where nbr is a wall
// Each thread owns V consecutive x cells of one (y) row of the flattened
// z-plane and marches up a chunk of z: the z neighbours stay in registers,
// x neighbours come from the vector and from adjacent lanes, y neighbours are
// the rows +-Nx of the same plane (cache hits from the neighbouring waves).
// ADJ: the reverse sweep applies the same operator to the adjoint field;
// hard-source cells hold prescribed values, so their adjoint is zero.
template <typename T, int V, bool ADJ>
__global__ __launch_bounds__(kBlock) void k_dense(const DenseParams p)
{
    using S = Store<T>;
    using C = typename S::C;
    using VT = Vec<T, V>;
    using VB = Vec<uint8_t, V>;
    const int lane = threadIdx.x & 63;

    // XCD-aware order: the hardware places block b on XCD b % 8, so give each
    // XCD a contiguous run of (z-chunk, plane-chunk) tiles and the y/z halo
    // rows of neighbouring tiles meet in the same L2.
    const unsigned nbk = gridDim.x, bid = blockIdx.x;
    const unsigned logical = (nbk % 8u == 0u) ? (bid % 8u) * (nbk / 8u) + bid / 8u : bid;
    const int pchunk = (int)(logical % (unsigned)p.nchunks_plane);
    const int zchunk = (int)(logical / (unsigned)p.nchunks_plane);

    const long long v = (long long)pchunk * kBlock + threadIdx.x;
    const bool active = v < p.nvec_plane;
    const long long cell0 = active ? v * V : 0;
    const int y = (int)(cell0 / p.nx);
    const int x0 = (int)(cell0 - (long long)y * p.nx);
    const int z0 = zchunk * p.zc;
    const int z1 = min(z0 + p.zc, p.nz);

    const long long n = p.ctl->n;
    C hard = (C)0;
    if (!ADJ) {
        if (p.hard_ref) hard = n < 10 ? (C)10 : (C)0;
        else if (n >= 0 && n < p.hard_len) hard = (C)p.hard_sig[n];
    }

    C sP = (C)1, sQ = (C)1, sR = (C)1;
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
            eR = min(max(14 - k2, -100), 100);
        }
        sP = exp2i(-p.ctl->exp[ip]);
        sQ = exp2i(-p.ctl->exp[iq]);
        sR = exp2i(eR);
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
    const int nx = p.nx;
    const C c2dt2 = (C)p.c2dt2, ih2 = (C)p.ih2;

    C pm[V], pc[V], pp[V];
    auto loadP = [&](long long idx, C *out) {
        const VT t = *reinterpret_cast<const VT *>(P + idx);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            if constexpr (S::kScaled) out[j] = S::ld(t.v[j]) * sP;
            else out[j] = S::ld(t.v[j]);
        }
    };
#pragma unroll
    for (int j = 0; j < V; ++j) pm[j] = pc[j] = (C)0;
    if (active && z0 >= 1) loadP((long long)(z0 - 1) * sxy + cell0, pm);
    if (active && z0 < p.nz) loadP((long long)z0 * sxy + cell0, pc);

    float lmax = 0.f;
    for (int z = z0; z < z1; ++z) {
        const long long base = (long long)z * sxy + cell0;
        C qv[V], yn[V], ys[V];
        uint8_t cd[V];
#pragma unroll
        for (int j = 0; j < V; ++j) pp[j] = qv[j] = yn[j] = ys[j] = (C)0, cd[j] = kTypeZero << 6;
        if (active) {
            if (z + 1 < p.nz) loadP(base + sxy, pp);
            const VT tq = *reinterpret_cast<const VT *>(Q + base);
            const VB tc = *reinterpret_cast<const VB *>(p.code + base);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                if constexpr (S::kScaled) qv[j] = S::ld(tq.v[j]) * sQ;
                else qv[j] = S::ld(tq.v[j]);
                cd[j] = tc.v[j];
            }
            if (y + 1 < p.ny) loadP(base + nx, yn);
            if (y > 0) loadP(base - nx, ys);
        }
        C left = __shfl_up(pc[V - 1], 1);
        C right = __shfl_down(pc[0], 1);
        if (active && lane == 0 && x0 > 0) {
            if constexpr (S::kScaled) left = S::ld(P[base - 1]) * sP;
            else left = S::ld(P[base - 1]);
        }
        if (active && lane == 63 && x0 + V < nx) {
            if constexpr (S::kScaled) right = S::ld(P[base + V]) * sP;
            else right = S::ld(P[base + V]);
        }
        VT out;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const uint8_t c = cd[j];
            const int type = c >> 6;
            C o = (C)0;
            if (type == kTypeUpdate) {
                const C pcv = pc[j];
                const C xp = (c & 1) ? pcv : (j + 1 < V ? pc[(j + 1) % V] : right);
                const C xm = (c & 2) ? pcv : (j > 0 ? pc[(j + V - 1) % V] : left);
                const C yp = (c & 4) ? pcv : yn[j];
                const C ym = (c & 8) ? pcv : ys[j];
                const C zp = (c & 16) ? pcv : pp[j];
                const C zm = (c & 32) ? pcv : pm[j];
                C sum = xp + xm;
                sum = sum + yp;
                sum = sum + ym;
                sum = sum + zp;
                sum = sum + zm;
                const C lap = (sum - (C)6 * pcv) * ih2;
                C a = (C)2, b = (C)1, g = (C)1;
                if (p.sp.w > 0) {
                    const int d = edge_dist(x0 + j, y, z, nx, p.ny, p.nz);
                    if (d < p.sp.w) {
                        a = (C)p.sp.a[d];
                        b = (C)p.sp.b[d];
                        g = (C)p.sp.g[d];
                    }
                }
                o = (a * pcv - b * qv[j]) + g * (c2dt2 * lap);
            } else if (type == kTypeHard) {
                o = hard;
            }
            if constexpr (S::kScaled) {
                lmax = fmaxf(lmax, fabsf((float)o));
                out.v[j] = S::st(o * sR);
            } else {
                out.v[j] = S::st(o);
            }
        }
        if (active) *reinterpret_cast<VT *>(R + base) = out;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            pm[j] = pc[j];
            pc[j] = pp[j];
        }
    }

    if constexpr (S::kScaled) {
        __shared__ float red[kBlock / 64];
        lmax = wave_max(lmax);
        if (lane == 0) red[threadIdx.x >> 6] = lmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = red[0];
            for (int w = 1; w < kBlock / 64; ++w) m = fmaxf(m, red[w]);
            atomicMax(&p.ctl->maxabs[n & 3][bid % kSlots], __float_as_uint(m));
        }
    }
}

struct TailParams {
    const void *P;
    const void *Q;
    void *R;
    long long nb;
    const uint32_t *bidx;
    const uint8_t *bmat;     // may be null: row 0
    double *s0, *s1, *s2, *s3; // forward: x1 x2 y1 y2; reverse: ybar1 ybar2 dbar1 dbar2
    const double *coeffs;    // [nmat][5]
    int nmat;
    float kconst;            // (c*dt)/h*0.95f
    StepCtl *ctl;
    int nsoft;
    const long long *soft_idx, *soft_bpos;
    const double *soft_sig;
    long long soft_len;
    int nrcv;
    const long long *rcv_idx, *rcv_bpos;
    double *ir;              // forward: [nrcv][ir_cap] output
    const double *gsig;      // reverse: dL/dIR [nrcv][ir_cap]
    long long ir_cap;
    double *trace_d, *trace_y; // [slots][nb]; slot = n - trace_base + 2
    long long trace_base, trace_slots;
    double *grad_part;       // reverse: [gridDim][nmat][5]
    int nx, ny, nz;
    Sponge sp;
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

__device__ inline void advance_step(StepCtl *ctl, long long n_next)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(&ctl->done, 1u);
        if (prev == gridDim.x - 1) {
            atomicExch(&ctl->done, 0u);
            atomicExch((unsigned long long *)&ctl->n, (unsigned long long)n_next);
            __threadfence();
        }
    }
}

// Boundary biquad over the compact list (simulator_gpu.cu:105-150), in order:
//   d = P[b] - Q[b]
//   y = b0 d + b1 x1 + b2 x2 - a1 y1 - a2 y2          (FP64 state)
//   R[b] = R[b] - K (0.1 y + 0.9 d) 0.4f               (evaluated in double)
// plus soft sources (added to R before the correction), receiver gather
// (after it) and the step counter.  The launch is ordered after k_dense on
// the stream, as the reference's two launches are.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tail(const TailParams p)
{
    using S = Store<T>;
    using C = typename S::C;
    const long long n = p.ctl->n;
    C sP = (C)1, sQ = (C)1, sR = (C)1, sRi = (C)1;
    if constexpr (S::kScaled) {
        sP = exp2i(-p.ctl->exp[(n - 1) & 3]);
        sQ = exp2i(-p.ctl->exp[(n - 2) & 3]);
        const int eR = p.ctl->exp[n & 3];
        sR = exp2i(eR);
        sRi = exp2i(-eR);
    }
    const T *__restrict__ P = (const T *)p.P;
    const T *__restrict__ Q = (const T *)p.Q;
    T *__restrict__ R = (T *)p.R;
    const bool in_sig = n >= 0 && n < p.soft_len;
    const bool in_ir = n >= 0 && n < p.ir_cap;
    float lmax = 0.f;

    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (e < p.nb) {
        const long long b = p.bidx[e];
        const int m = p.bmat ? p.bmat[e] : 0;
        const double *cf = p.coeffs + 5 * m;
        C pv = S::ld(P[b]), qv = S::ld(Q[b]);
        if constexpr (S::kScaled) pv = pv * sP, qv = qv * sQ;
        const C dif = pv - qv;
        const double in = (double)dif;
        const double x1 = p.s0[e], x2 = p.s1[e], y1 = p.s2[e], y2 = p.s3[e];
        const double yv = cf[0] * in + cf[1] * x1 + cf[2] * x2 - cf[3] * y1 - cf[4] * y2;
        p.s1[e] = x1;
        p.s0[e] = in;
        p.s3[e] = y1;
        p.s2[e] = yv;
        C r = S::ld(R[b]);
        if constexpr (S::kScaled) r = r * sRi;
        for (int k = 0; k < p.nsoft; ++k)
            if (p.soft_bpos[k] == e && in_sig) r = r + (C)p.soft_sig[(long long)k * p.soft_len + n];
        r = (C)((double)r - (double)p.kconst * (yv * 0.1 + in * 0.9) * (double)0.4f);
        if constexpr (S::kScaled) {
            R[b] = S::st(r * sR);
            lmax = fabsf((float)r);
        } else {
            R[b] = S::st(r);
        }
        if (p.trace_d) {
            const long long slot = n - p.trace_base + 2;
            if (slot >= 0 && slot < p.trace_slots) {
                p.trace_d[slot * p.nb + e] = in;
                p.trace_y[slot * p.nb + e] = yv;
            }
        }
        if (in_ir)
            for (int k = 0; k < p.nrcv; ++k)
                if (p.rcv_bpos[k] == e) {
                    C rv = r;
                    if constexpr (S::kScaled) rv = S::ld(S::st(r * sR)) * sRi; // the stored value
                    p.ir[(long long)k * p.ir_cap + n] = (double)rv;
                }
    }
    if (blockIdx.x == 0) {
        const int k = threadIdx.x;
        if (k < p.nsoft && p.soft_bpos[k] < 0 && in_sig) {
            const long long i = p.soft_idx[k];
            C r = S::ld(R[i]);
            if constexpr (S::kScaled) r = r * sRi;
            r = r + (C)p.soft_sig[(long long)k * p.soft_len + n];
            if constexpr (S::kScaled) {
                R[i] = S::st(r * sR);
                lmax = fmaxf(lmax, fabsf((float)r));
            } else {
                R[i] = S::st(r);
            }
        }
        __syncthreads();
        if (k < p.nrcv && p.rcv_bpos[k] < 0 && in_ir) {
            C r = S::ld(R[p.rcv_idx[k]]);
            if constexpr (S::kScaled) r = r * sRi;
            p.ir[(long long)k * p.ir_cap + n] = (double)r;
        }
    }
    if constexpr (S::kScaled) {
        __shared__ float red[kBlock / 64];
        lmax = wave_max(lmax);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float mx = red[0];
            for (int w = 1; w < kBlock / 64; ++w) mx = fmaxf(mx, red[w]);
            atomicMax(&p.ctl->maxabs[n & 3][blockIdx.x % kSlots], __float_as_uint(mx));
        }
    }
    advance_step(p.ctl, n + 1);
}

// Reverse boundary step.  With w = g*rbar the adjoint of the field after
// k_dense<ADJ> (g = sponge factor of the cell), for each entry:
//   rbar   = w/g + dbar1 - dbar2 (+ dL/dIR if a receiver sits here)
//   ybar   = -0.1 kap rbar - a1 ybar1 - a2 ybar2          kap = K*0.4f
//   dbar   = -0.9 kap rbar + b0 ybar + b1 ybar1 + b2 ybar2
//   dL/d(b0,b1,b2,a1,a2) += ybar * (d^n, d^{n-1}, d^{n-2}, -y^{n-1}, -y^{n-2})
// where d and y come from the forward trace of this segment.  State arrays:
// s0 ybar1, s1 ybar2, s2 dbar1, s3 dbar2.  Per-block gradient partials are
// summed in a fixed order, so the result is deterministic.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tail_adj(const TailParams p)
{
    using C = T;
    __shared__ double part[kBlock / 64][8 * 5];
    const long long n = p.ctl->n;
    const bool in_g = n >= 0 && n < p.ir_cap;
    T *__restrict__ R = (T *)p.R;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    double gv[5] = {0, 0, 0, 0, 0};
    int m = -1;
    if (e < p.nb) {
        const long long b = p.bidx[e];
        m = p.bmat ? p.bmat[e] : 0;
        const double *cf = p.coeffs + 5 * m;
        double g = 1.0;
        sponge_g(p, b, g);
        double rb = (double)R[b] / g;
        const double yb1 = p.s0[e], yb2 = p.s1[e], db1 = p.s2[e], db2 = p.s3[e];
        rb = rb + db1 - db2;
        if (in_g)
            for (int k = 0; k < p.nrcv; ++k)
                if (p.rcv_bpos[k] == e) rb += p.gsig[(long long)k * p.ir_cap + n];
        const double kap = (double)p.kconst * (double)0.4f;
        const double yb = -0.1 * kap * rb - cf[3] * yb1 - cf[4] * yb2;
        const double db = -0.9 * kap * rb + cf[0] * yb + cf[1] * yb1 + cf[2] * yb2;
        const long long sl = n - p.trace_base + 2;
        const double dn = p.trace_d[sl * p.nb + e];
        const double x1 = p.trace_d[(sl - 1) * p.nb + e];
        const double x2 = p.trace_d[(sl - 2) * p.nb + e];
        const double y1 = p.trace_y[(sl - 1) * p.nb + e];
        const double y2 = p.trace_y[(sl - 2) * p.nb + e];
        gv[0] = yb * dn;
        gv[1] = yb * x1;
        gv[2] = yb * x2;
        gv[3] = -yb * y1;
        gv[4] = -yb * y2;
        p.s1[e] = yb1;
        p.s0[e] = yb;
        p.s3[e] = db1;
        p.s2[e] = db;
        R[b] = (C)(rb * g);
    }
    if (blockIdx.x == 0) {
        const int k = threadIdx.x;
        if (k < p.nrcv && p.rcv_bpos[k] < 0 && in_g) {
            const long long i = p.rcv_idx[k];
            double g = 1.0;
            sponge_g(p, i, g);
            R[i] = (C)((double)R[i] + g * p.gsig[(long long)k * p.ir_cap + n]);
        }
    }
    // per-material block sums, fixed reduction order
    for (int mm = 0; mm < p.nmat; ++mm) {
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double s = wave_sum(m == mm ? gv[c] : 0.0);
            if (lane == 0) part[wave][mm * 5 + c] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < 5 * p.nmat) {
        double s = part[0][threadIdx.x];
        for (int w = 1; w < kBlock / 64; ++w) s = s + part[w][threadIdx.x];
        double *dst = p.grad_part + (long long)blockIdx.x * 5 * p.nmat + threadIdx.x;
        *dst = *dst + s;
    }
    advance_step(p.ctl, n - 1);
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

// Sets the step counter (and, with reset, the whole control block) without a
// memset, so it can sit inside a captured graph.
__global__ void k_ctl_set(StepCtl *ctl, long long n, int reset)
{
    const int t = threadIdx.x;
    if (reset) {
        if (t < 4) {
            ctl->exp[t] = 0;
            for (int s = 0; s < kSlots; ++s) ctl->maxabs[t][s] = 0u;
        }
        if (t == 0) {
            ctl->done = 0u;
            ctl->err = 0;
        }
    }
    if (t == 0) ctl->n = n;
}

__global__ void k_fill_f64(double *p, long long n, double v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Copies the filter state at a segment start into trace slots 0 (step s0-2)
// and 1 (step s0-1): x2,y2 and x1,y1.
__global__ void k_trace_seed(const double *x1, const double *x2, const double *y1, const double *y2,
                             double *trace_d, double *trace_y, long long nb)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nb;
         e += (long long)gridDim.x * blockDim.x) {
        trace_d[e] = x2[e];
        trace_y[e] = y2[e];
        trace_d[nb + e] = x1[e];
        trace_y[nb + e] = y1[e];
    }
}

} // namespace virs
