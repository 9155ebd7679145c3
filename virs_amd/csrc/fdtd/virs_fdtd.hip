// libvirs_hip.so: the C ABI of include/virs_fdtd.h on top of kernels.hpp.
//
// Host orchestration of the step loop (the reference's fdtd_gpu_step,
// src/simulator/simulator_gpu.cu:184-332), without its per-step
// cudaMemcpyToSymbol, three device synchronisations and 4-byte receiver
// readback: a step is two stream-ordered launches (stencil, then boundary +
// points + step counter), receivers accumulate in a device buffer, and
// batches of steps can be replayed as HIP graphs.
#include "kernels.hpp"
#include "virs_fdtd.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace virs;

constexpr int kSnapSlots = 4; // layer snapshots outstanding at once

// Device staging of virs_plane_read_noalloc: one XZ plane of at most
// kStaticPlaneCells floats (8 MiB), held by every GPU the library runs on
// until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is a stand-in, not a device address): virs_create
// resolves each device's copy once with hipGetSymbolAddress, with that device
// current, and the entry point uses that address.
constexpr long long kStaticPlaneCells = 2LL * 1024 * 1024;
static __device__ __align__(256) float scratch[kStaticPlaneCells];
constexpr int kMaxDevices = 64;

struct virs_ctx {
    virs_desc d{};
    hipStream_t stream = nullptr;
    long long nx = 0, ny = 0, nz = 0, N = 0;
    int vw = 1;            // cells per thread along x in the stencil
    size_t esz = 4;        // bytes per stored pressure
    float dt = 0, c2dt2 = 0, ih2 = 0, kconst = 0;
    Sponge sp{};

    void *f[3] = {nullptr, nullptr, nullptr};
    int ip = 0, iq = 1, ir = 2; // rotation: P, Q, R
    uint8_t *code = nullptr;
    uint8_t *normals = nullptr;
    bool have_flags = false;
    long long n_hard = 0;          // interior flag-3 cells (hard source)
    std::vector<uint32_t> h_flag2; // interior flag-2 cells of the uploaded flags
    int nchunks_plane = 0;         // 256-thread chunks per z-plane of the stencil
    // [nchunks_plane][4 waves][nz] {first entry, count}: the list slice of
    // each wave's cells per plane (DenseParams::wtab)
    uint2 *wtab = nullptr;
    bool fused = false;            // boundary filter runs inside the stencil pass
    long long *neg = nullptr;      // kMaxPoints x -1: point positions when fused
    int cus = 0;

    long long nb = 0;
    uint32_t *bidx = nullptr;
    uint8_t *bmat = nullptr;
    // filter state (BoundaryState).  FP32/FP64 fields: st[0..1] the history
    // of the filter inputs d in the compute type (xs), st[2..3] of the
    // outputs y in FP64 (ys), which the adjoint's recording writes before
    // each snapshot, and st[4..5] the direct-form-II state ring w in FP64
    // (ws).  16-bit fields: st[0..1] the transposed-direct-form-II states
    // s1, s2 in FP32 (ts).
    void *st[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void *hv[2] = {nullptr, nullptr}; // list-cell pressures by step parity (TailParams::hv)
    std::vector<uint32_t> h_bidx;
    std::vector<uint8_t> h_bmat;

    double *coeffs = nullptr;
    int nmat = 0;

    StepCtl *ctl = nullptr;
    StepCtl *ctl_adj = nullptr;

    double *hard_sig = nullptr;
    long long hard_len = 0;
    int hard_ref = 1;

    int nsoft = 0;
    long long soft_len = 0;
    long long *soft_idx = nullptr, *soft_bpos = nullptr;
    double *soft_sig = nullptr;
    std::vector<long long> h_soft_idx;
    std::vector<double> h_soft_sig;

    int nrcv = 0;
    long long *rcv_idx = nullptr, *rcv_bpos = nullptr;
    // TailParams::rcv_src/rcv_next/rcv_first/src_rcv, [4][kMaxPoints]
    int *rcv_aux = nullptr;
    bool points_on_list = false; // some soft source or receiver is a list cell
    std::vector<long long> h_rcv_idx;
    double *ir_buf = nullptr;
    long long ir_cap = 0;

    long long steps = 0;   // steps run since virs_start_simulation
    long long counted = 0; // steps reported (early stop truncates)
    bool stopped = false;
    int es_enable = 0;
    float es_target = 0.0009f, percent = 0.f;

    // graphs by starting rotation phase: kGraphSteps steps each, and one
    // shorter graph per phase for the remainder of a batch (tail_steps[k]
    // steps)
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t tail_graph[3] = {nullptr, nullptr, nullptr};
    int tail_steps[3] = {0, 0, 0};
    // when set, the next stencil or list launch records its own start and
    // stop in these events (hipExtLaunchKernel: the kernel's execution, not
    // the dispatch around it); launch_step sets them for sampled steps
    hipEvent_t kev[2] = {nullptr, nullptr};

    // adjoint
    int ncp = 0;
    // steps of the run recorded by virs_forward_record; 0 when there is none
    // or when a later call (materials, flags, sources, boundary list) changed
    // what the recorded run computed
    long long rec_steps = 0, seg = 0;
    // set by virs_adjoint when its recomputation left the fields, filter
    // state and step counter mid-run: stepping and field reads are refused
    // until virs_start_simulation or virs_forward_record resets the run
    bool restart_needed = false;
    std::vector<void *> cp_field;   // [ncp][2] stored fields
    std::vector<uint8_t *> cp_state; // [ncp] the four state rings back to back
    // two boundary-trace buffers (segment k uses buffer k & 1), elements of
    // the field's compute type; a segment's recomputation on stream2 fills
    // one while the reverse sweep of the next segment reads the other
    void *trace_d[2] = {nullptr, nullptr}, *trace_y[2] = {nullptr, nullptr};
    long long trace_slots = 0;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_rec[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr};
    void *af[3] = {nullptr, nullptr, nullptr};
    void *ast[4] = {nullptr, nullptr, nullptr, nullptr}; // reverse rings: dbar x2, ybar x2 (FP64)
    double *gsig = nullptr;
    double *grad_part = nullptr;
    long long grad_blocks = 0;

    // layer snapshots: kSnapSlots gathered XZ planes on the device, their
    // pinned host copies, the stream the copies run on (so they overlap the
    // next steps), gather and copy events; snap_ticket[k] is the ticket held
    // by slot k (-1 when free), tickets are handed out in order
    float *snap_dev = nullptr, *snap_host = nullptr;
    hipStream_t snap_stream = nullptr;
    hipEvent_t snap_gather[kSnapSlots] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t snap_ev[kSnapSlots] = {nullptr, nullptr, nullptr, nullptr};
    int snap_ticket[kSnapSlots] = {-1, -1, -1, -1};
    int snap_next = 0;

    float *scratch_dev = nullptr; // this device's address of `scratch`

    std::string err;
};

namespace {

int fail(virs_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPC(ctx, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), VIRS_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                  \
                        hipGetErrorString(e_), __FILE__, __LINE__);                            \
    } while (0)

template <typename P>
int dalloc(virs_ctx *c, P **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, VIRS_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return VIRS_OK;
}

template <typename P>
void dfree(P *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// The device address of `scratch` on the current device (ordinal dev),
// resolved once per device and kept.
int static_scratch(virs_ctx *c, int dev, float **out)
{
    static std::mutex mu;
    static float *addr[kMaxDevices] = {};
    if (dev < 0 || dev >= kMaxDevices) return fail(c, VIRS_ERR_ARG, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(c, VIRS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        addr[dev] = (float *)p;
    }
    *out = addr[dev];
    return VIRS_OK;
}

int grid_for(long long n) { return (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), 65536); }

// Steps per replayed graph; a multiple of 3 keeps the buffer rotation phase.
constexpr int kGraphSteps = 30;

void drop_graphs(virs_ctx *c)
{
    for (int k = 0; k < 3; ++k) {
        if (c->graph[k]) (void)hipGraphExecDestroy(c->graph[k]);
        if (c->tail_graph[k]) (void)hipGraphExecDestroy(c->tail_graph[k]);
        c->graph[k] = c->tail_graph[k] = nullptr;
        c->tail_steps[k] = 0;
    }
}

// Position of linear index `idx` in the sorted boundary list, or -1.
long long bpos_of(const virs_ctx *c, long long idx)
{
    auto it = std::lower_bound(c->h_bidx.begin(), c->h_bidx.end(), (uint32_t)idx);
    if (idx <= (long long)UINT32_MAX && it != c->h_bidx.end() && (long long)*it == idx)
        return (long long)(it - c->h_bidx.begin());
    return -1;
}

int upload_points(virs_ctx *c)
{
    std::vector<long long> sb(c->nsoft), rb(c->nrcv);
    c->points_on_list = false;
    for (int k = 0; k < c->nsoft; ++k) {
        sb[k] = bpos_of(c, c->h_soft_idx[k]);
        c->points_on_list |= sb[k] >= 0;
    }
    for (int k = 0; k < c->nrcv; ++k) {
        rb[k] = bpos_of(c, c->h_rcv_idx[k]);
        c->points_on_list |= rb[k] >= 0;
    }
    if (c->nsoft) {
        HIPC(c, hipMemcpy(c->soft_idx, c->h_soft_idx.data(), sizeof(long long) * c->nsoft, hipMemcpyHostToDevice));
        HIPC(c, hipMemcpy(c->soft_bpos, sb.data(), sizeof(long long) * c->nsoft, hipMemcpyHostToDevice));
    }
    if (c->nrcv) {
        HIPC(c, hipMemcpy(c->rcv_idx, c->h_rcv_idx.data(), sizeof(long long) * c->nrcv, hipMemcpyHostToDevice));
        HIPC(c, hipMemcpy(c->rcv_bpos, rb.data(), sizeof(long long) * c->nrcv, hipMemcpyHostToDevice));
    }
    // the soft source on each receiver's cell, the receivers that share a
    // cell as chains in index order, and each source's first receiver
    std::vector<int> aux(4 * (size_t)kMaxPoints, -1);
    for (int k = 0; k < c->nrcv; ++k) {
        for (int j = 0; j < c->nsoft; ++j)
            if (c->h_soft_idx[j] == c->h_rcv_idx[k]) aux[k] = j;
        for (int j = k + 1; j < c->nrcv && aux[kMaxPoints + k] < 0; ++j)
            if (c->h_rcv_idx[j] == c->h_rcv_idx[k]) aux[kMaxPoints + k] = j;
        aux[2 * kMaxPoints + k] = 1;
        for (int j = 0; j < k; ++j)
            if (c->h_rcv_idx[j] == c->h_rcv_idx[k]) aux[2 * kMaxPoints + k] = 0;
    }
    for (int j = 0; j < c->nsoft; ++j)
        for (int k = c->nrcv - 1; k >= 0; --k)
            if (c->h_rcv_idx[k] == c->h_soft_idx[j]) aux[3 * kMaxPoints + j] = k;
    HIPC(c, hipMemcpy(c->rcv_aux, aux.data(), sizeof(int) * aux.size(), hipMemcpyHostToDevice));
    return VIRS_OK;
}

bool is16(const virs_ctx *c);

// The forward filter state of a step: the transposed-direct-form-II states
// for 16-bit fields; the direct-form-II ring for FP32/FP64 fields, with the
// d/y history rings only when `history` is set (BoundaryState).
BoundaryState boundary_state(const virs_ctx *c, bool history = false)
{
    BoundaryState b{};
    b.nb = c->nb;
    b.bmat = c->bmat;
    if (is16(c)) {
        b.ts[0] = c->st[0];
        b.ts[1] = c->st[1];
    } else {
        b.ws[0] = c->st[4];
        b.ws[1] = c->st[5];
        if (history) {
            b.xs[0] = c->st[0];
            b.xs[1] = c->st[1];
            b.ys[0] = c->st[2];
            b.ys[1] = c->st[3];
        }
    }
    b.coeffs = c->coeffs;
    b.nmat = c->nmat;
    b.kconst = c->kconst;
    return b;
}

DenseParams dense_params(const virs_ctx *c, const void *P, const void *Q, void *R, StepCtl *ctl,
                         int slot)
{
    DenseParams p{};
    p.P = P;
    p.Q = Q;
    p.R = R;
    p.code = c->code;
    p.nx = (int)c->nx;
    p.ny = (int)c->ny;
    p.nz = (int)c->nz;
    p.sxy = c->nx * c->ny;
    p.nvec_plane = p.sxy / c->vw;
    p.nchunks_plane = c->nchunks_plane;
    p.c2dt2 = c->c2dt2;
    p.ih2 = c->ih2;
    p.ctl = ctl;
    p.slot = slot;
    p.hard_sig = c->hard_sig;
    p.hard_len = c->hard_len;
    p.hard_ref = c->hard_ref;
    p.nsoft = c->nsoft;
    p.soft_sig = c->soft_sig;
    p.soft_len = c->soft_len;
    p.sp = c->sp;
    p.sr = c->d.rounding == VIRS_ROUND_STOCHASTIC;
    p.wtab = nullptr;
    p.bs = boundary_state(c);
    return p;
}

TailParams tail_params(const virs_ctx *c, const void *P, const void *Q, void *R, StepCtl *ctl,
                       int slot, int next_slot)
{
    TailParams t{};
    t.P = P;
    t.Q = Q;
    t.R = R;
    t.bidx = c->bidx;
    t.bs = boundary_state(c);
    t.ctl = ctl;
    t.slot = slot;
    t.next_slot = next_slot;
    t.nsoft = c->nsoft;
    t.soft_idx = c->soft_idx;
    t.soft_bpos = c->soft_bpos;
    t.soft_sig = c->soft_sig;
    t.soft_len = c->soft_len;
    t.nrcv = c->nrcv;
    t.rcv_idx = c->rcv_idx;
    t.rcv_bpos = c->rcv_bpos;
    t.rcv_src = c->rcv_aux;
    t.rcv_next = c->rcv_aux + kMaxPoints;
    t.rcv_first = c->rcv_aux + 2 * kMaxPoints;
    t.src_rcv = c->rcv_aux + 3 * kMaxPoints;
    t.list_points = c->points_on_list ? 1 : 0;
    t.ir = c->ir_buf;
    t.ir_cap = c->ir_cap;
    t.nx = (int)c->nx;
    t.ny = (int)c->ny;
    t.nz = (int)c->nz;
    t.sp = c->sp;
    t.sr = c->d.rounding == VIRS_ROUND_STOCHASTIC;
    t.hv[0] = c->fused ? nullptr : c->hv[0];
    t.hv[1] = c->fused ? nullptr : c->hv[1];
    return t;
}

// List launches: block 0 takes the points and the step counter, blocks
// 1.. one list entry per thread.
long long list_blocks(long long nb, int per_thread = 1, int block = kBlock)
{
    return (nb + (long long)block * per_thread - 1) / ((long long)block * per_thread);
}
int tail_grid(long long nb, int per_thread = 1, int block = kBlock) { return (int)(1 + list_blocks(nb, per_thread, block)); }

// Stencil grid.  capacity = resident blocks of the whole GPU: occupancy x
// CUs, the occupancy capped at VIRS_DENSE_OCC_CAP blocks (12 waves) per CU.
// Fewer blocks that each march more planes beat more resident waves here:
// on the 512x384x256 room, planning for 2 / 3 / 4 / 5 / 6 blocks per CU
// gave FP32 steps of 363 / 399 / 377 / 373 / 390 Gvox/s and FP16 steps of
// 557 / 648 / 626 / 626 Gvox/s; with the current 16-bit stencil, 2 / 3 / 4
// blocks give FP16 613 / 710 / 666 and FP32 364 / 399 / 376.  Tile mode (one block per chunk x z-range)
// is used when whole z-ranges of every chunk fill at least 85% of the
// capacity in one round; otherwise the (chunk, z) planes are split into
// `capacity` equal ranges.  A descriptor value dense_blocks > 0 forces the
// split mode with that grid.
#ifndef VIRS_DENSE_OCC_CAP
#define VIRS_DENSE_OCC_CAP 3 // blocks per CU the stencil grid plans for at most (0 = no bound)
#endif
template <typename K>
int dense_grid(const virs_ctx *c, K kernel, int &zranges)
{
    zranges = 0;
    const long long W = (long long)c->nchunks_plane * c->nz;
    if (c->d.dense_blocks > 0) return (int)std::min<long long>(c->d.dense_blocks, W);
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kBlock, 0) != hipSuccess || occ <= 0)
        occ = 4;
    if (VIRS_DENSE_OCC_CAP > 0) occ = std::min(occ, VIRS_DENSE_OCC_CAP);
    const long long cap = (long long)occ * std::max(c->cus, 1);
    const long long zr = std::min<long long>(cap / std::max(c->nchunks_plane, 1), c->nz);
    if (zr >= 1 && zr * c->nchunks_plane * 100 >= cap * 85) {
        zranges = (int)zr;
        return (int)(zr * c->nchunks_plane);
    }
    long long g = cap;
    if (g >= 8) g -= g % 8;
    return (int)std::max<long long>(1, std::min<long long>(g, W));
}

template <typename T, int V, bool ADJ, bool FUSED, bool SPONGE, bool SR, bool HARD, int LY>
hipError_t launch_dense_kl(const virs_ctx *c, const DenseParams &p)
{
    auto k = k_dense<T, V, ADJ, FUSED, SPONGE, SR, HARD, LY>;
    DenseParams q = p;
    const int grid = dense_grid(c, k, q.zranges);
    if (c->kev[0])
        hipExtLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, c->stream, c->kev[0], c->kev[1], 0, q);
    else
        hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, c->stream, q);
    return hipGetLastError();
}

// The stencil stages its planes through LDS when a row has at most
// 2 * kHaloVecs vectors (k_dense LY), except with the fused boundary.
template <typename T, int V, bool ADJ, bool FUSED, bool SPONGE, bool SR, bool HARD = true>
hipError_t launch_dense_k(const virs_ctx *c, const DenseParams &p)
{
    if constexpr (!FUSED) {
        const int h = c->nx / V;
        if (h <= kHaloVecs) return launch_dense_kl<T, V, ADJ, FUSED, SPONGE, SR, HARD, 1>(c, p);
        if (h <= 2 * kHaloVecs) return launch_dense_kl<T, V, ADJ, FUSED, SPONGE, SR, HARD, 2>(c, p);
    }
    return launch_dense_kl<T, V, ADJ, FUSED, SPONGE, SR, HARD, 0>(c, p);
}

template <typename T, int V, bool HARD>
hipError_t launch_dense_s16(const virs_ctx *c, const DenseParams &p, bool sponge)
{
    if (p.sr) {
        if (p.wtab)
            return sponge ? launch_dense_k<T, V, false, true, true, true, HARD>(c, p)
                          : launch_dense_k<T, V, false, true, false, true, HARD>(c, p);
        return sponge ? launch_dense_k<T, V, false, false, true, true, HARD>(c, p)
                      : launch_dense_k<T, V, false, false, false, true, HARD>(c, p);
    }
    if (p.wtab)
        return sponge ? launch_dense_k<T, V, false, true, true, false, HARD>(c, p)
                      : launch_dense_k<T, V, false, true, false, false, HARD>(c, p);
    return sponge ? launch_dense_k<T, V, false, false, true, false, HARD>(c, p)
                  : launch_dense_k<T, V, false, false, false, false, HARD>(c, p);
}

template <typename T, int V, bool ADJ>
hipError_t launch_dense_v(const virs_ctx *c, const DenseParams &p)
{
    if constexpr (sizeof(T) * V > 16) {
        return hipErrorInvalidValue;
    } else {
        const bool sponge = p.sp.w > 0;
        if constexpr (Store<T>::kScaled) {
            // stochastic or nearest storage rounding and the presence of
            // hard cells, chosen at compile time
            if (c->n_hard > 0) return launch_dense_s16<T, V, true>(c, p, sponge);
            return launch_dense_s16<T, V, false>(c, p, sponge);
        } else {
            if constexpr (!ADJ) {
                if (p.wtab) {
                    return sponge ? launch_dense_k<T, V, false, true, true, false>(c, p)
                                  : launch_dense_k<T, V, false, true, false, false>(c, p);
                }
            }
            return sponge ? launch_dense_k<T, V, ADJ, false, true, false>(c, p)
                          : launch_dense_k<T, V, ADJ, false, false, false>(c, p);
        }
    }
}

template <typename T, bool ADJ>
hipError_t launch_dense_t(const virs_ctx *c, const DenseParams &p)
{
    switch (c->vw) {
    case 8: return launch_dense_v<T, 8, ADJ>(c, p);
    case 4: return launch_dense_v<T, 4, ADJ>(c, p);
    case 2: return launch_dense_v<T, 2, ADJ>(c, p);
    default: return launch_dense_v<T, 1, ADJ>(c, p);
    }
}

hipError_t launch_dense(const virs_ctx *c, const DenseParams &p, bool adj)
{
    switch (c->d.precision) {
    case VIRS_FP16: return launch_dense_t<f16, false>(c, p);
    case VIRS_BF16: return launch_dense_t<bf16, false>(c, p);
    case VIRS_FP64: return adj ? launch_dense_t<double, true>(c, p) : launch_dense_t<double, false>(c, p);
    default: return adj ? launch_dense_t<float, true>(c, p) : launch_dense_t<float, false>(c, p);
    }
}

template <typename T>
void launch_tail_t(const virs_ctx *c, const TailParams &t)
{
    const int g = tail_grid(t.bs.nb, kTailEntries, kTailBlock);
    if (c->kev[0])
        hipExtLaunchKernelGGL(k_tail<T>, dim3(g), dim3(kTailBlock), 0, c->stream, c->kev[0], c->kev[1], 0, t);
    else
        hipLaunchKernelGGL(k_tail<T>, dim3(g), dim3(kTailBlock), 0, c->stream, t);
}

hipError_t launch_tail(const virs_ctx *c, const TailParams &t)
{
    switch (c->d.precision) {
    case VIRS_FP16: launch_tail_t<f16>(c, t); break;
    case VIRS_BF16: launch_tail_t<bf16>(c, t); break;
    case VIRS_FP64: launch_tail_t<double>(c, t); break;
    default: launch_tail_t<float>(c, t); break;
    }
    return hipGetLastError();
}

// Seeds TailParams::hv from the current fields for the next step n.
int stash_boundary(virs_ctx *c, long long n)
{
    if (!c->nb || !c->hv[0]) return VIRS_OK;
    const void *P = c->f[c->ip], *Q = c->f[c->iq];
    void *pn = c->hv[n & 1], *pq = c->hv[(n + 1) & 1];
    const int g = grid_for(c->nb);
    switch (c->d.precision) {
    case VIRS_FP16:
        hipLaunchKernelGGL(k_stash<f16>, dim3(g), dim3(256), 0, c->stream, (const f16 *)P, (const f16 *)Q, c->bidx,
                           (f16 *)pn, (f16 *)pq, c->nb);
        break;
    case VIRS_BF16:
        hipLaunchKernelGGL(k_stash<bf16>, dim3(g), dim3(256), 0, c->stream, (const bf16 *)P, (const bf16 *)Q,
                           c->bidx, (bf16 *)pn, (bf16 *)pq, c->nb);
        break;
    case VIRS_FP64:
        hipLaunchKernelGGL(k_stash<double>, dim3(g), dim3(256), 0, c->stream, (const double *)P,
                           (const double *)Q, c->bidx, (double *)pn, (double *)pq, c->nb);
        break;
    default:
        hipLaunchKernelGGL(k_stash<float>, dim3(g), dim3(256), 0, c->stream, (const float *)P, (const float *)Q,
                           c->bidx, (float *)pn, (float *)pq, c->nb);
        break;
    }
    HIPC(c, hipGetLastError());
    return VIRS_OK;
}

hipError_t launch_tail_adj(const virs_ctx *c, const TailParams &t)
{
    const int g = tail_grid(t.bs.nb, 1, kTailAdjBlock);
    if (c->d.precision == VIRS_FP64)
        hipLaunchKernelGGL(k_tail_adj<double>, dim3(g), dim3(kTailAdjBlock), 0, c->stream, t);
    else
        hipLaunchKernelGGL(k_tail_adj<float>, dim3(g), dim3(kTailAdjBlock), 0, c->stream, t);
    return hipGetLastError();
}

// One forward step on the context's current rotation; optionally records the
// boundary trace for the adjoint, or (history) the d/y history rings that
// seed a snapshot's trace.
int launch_step(virs_ctx *c, void *trace_d = nullptr, void *trace_y = nullptr,
                long long trace_base = 0, bool receivers = true, hipEvent_t *ev = nullptr,
                bool history = false)
{
    DenseParams p = dense_params(c, c->f[c->ip], c->f[c->iq], c->f[c->ir], c->ctl, c->ip);
    TailParams t = tail_params(c, c->f[c->ip], c->f[c->iq], c->f[c->ir], c->ctl, c->ip, c->ir);
    BoundaryState &bs = c->fused ? p.bs : t.bs;
    if (history) bs = boundary_state(c, true);
    bs.trace_d = trace_d;
    bs.trace_y = trace_y;
    bs.trace_base = trace_base;
    bs.trace_slots = c->trace_slots;
    if (c->fused) {
        p.wtab = c->wtab;
        p.bidx = c->bidx;
        t.bs.nb = 0;
        t.soft_bpos = c->neg;
        t.rcv_bpos = c->neg;
    }
    if (!receivers) t.nrcv = 0;
    // ev (4 events, may be null): start and stop of the stencil launch, then
    // of the list launch
    if (ev) c->kev[0] = ev[0], c->kev[1] = ev[1];
    hipError_t e = launch_dense(c, p, false);
    if (ev) c->kev[0] = ev[2], c->kev[1] = ev[3];
    if (e == hipSuccess) e = launch_tail(c, t);
    c->kev[0] = c->kev[1] = nullptr;
    HIPC(c, e);
    const int ip = c->ip;
    c->ip = c->ir;
    c->ir = c->iq;
    c->iq = ip;
    return VIRS_OK;
}

// Builds the slice table that lets the stencil launch run the boundary
// filter of its own cells: for wave w of chunk ch at plane z, the list
// entries whose cells lie in the wave's 64*V cells of the plane
// ([(ch*256 + w*64)*V, +64*V) clipped to the plane) are the contiguous
// slice {first, count} of the ascending list.  One walk over the list.
int refresh_fusion(virs_ctx *c)
{
    dfree(c->wtab);
    c->fused = false;
    if (!c->d.fuse_boundary || c->nb == 0) return VIRS_OK;
    const long long nch = c->nchunks_plane, sxy = c->nx * c->ny, nz = c->nz;
    const long long wcells = 64LL * c->vw;
    std::vector<uint32_t> tab((size_t)(nch * kWaves * nz) * 2);
    size_t e = 0;
    const size_t nb = c->h_bidx.size();
    for (long long z = 0; z < nz; ++z)
        for (long long ch = 0; ch < nch; ++ch)
            for (long long w = 0; w < kWaves; ++w) {
                const long long lo = std::min(sxy, (ch * kWaves + w) * wcells), hi = std::min(sxy, lo + wcells);
                const long long glo = z * sxy + lo, ghi = z * sxy + hi;
                while (e < nb && (long long)c->h_bidx[e] < glo) ++e;
                size_t f = e;
                while (f < nb && (long long)c->h_bidx[f] < ghi) ++f;
                const size_t row = (size_t)((ch * kWaves + w) * nz + z);
                tab[2 * row] = (uint32_t)e;
                tab[2 * row + 1] = (uint32_t)(f - e);
                e = f;
            }
    int rc = dalloc(c, &c->wtab, sizeof(uint32_t) * tab.size());
    if (rc) return rc;
    HIPC(c, hipMemcpy(c->wtab, tab.data(), sizeof(uint32_t) * tab.size(), hipMemcpyHostToDevice));
    c->fused = true;
    return VIRS_OK;
}

int check_ready(virs_ctx *c)
{
    if (!c) return VIRS_ERR_ARG;
    if (!c->have_flags) return fail(c, VIRS_ERR_STATE, "flags not uploaded");
    if (!c->coeffs && c->nb > 0) return fail(c, VIRS_ERR_STATE, "materials not set");
    if (c->ir_cap == 0 && c->nrcv > 0) return fail(c, VIRS_ERR_STATE, "virs_start_simulation not called");
    HIPC(c, hipSetDevice(c->d.device));
    return VIRS_OK;
}

// bytes of the compute type C of a field (float, or double for FP64): the
// element of the filter-input rings and of the adjoint trace
size_t compute_bytes(const virs_ctx *c) { return c->d.precision == VIRS_FP64 ? 8 : 4; }
bool is16(const virs_ctx *c) { return c->d.precision == VIRS_FP16 || c->d.precision == VIRS_BF16; }
// cell codes: a byte per cell, half a byte in the 16-bit formats
size_t code_bytes(const virs_ctx *c) { return is16(c) ? (size_t)(c->N + 1) / 2 : (size_t)c->N; }

// bytes of filter-state array k (virs_ctx::st): FP32/FP64 fields: 0 and 1
// hold inputs in the compute type, 2 and 3 outputs in FP64, 4 and 5 the
// direct-form-II state in FP64; 16-bit fields: 0 and 1 the states s1, s2
// in FP32
size_t state_bytes(const virs_ctx *c, int k)
{
    const size_t el = is16(c) ? (k < 2 ? sizeof(float) : 0)
                              : (k < 2 ? compute_bytes(c) : sizeof(double));
    return el * (size_t)std::max<long long>(1, c->nb);
}
// filter-state arrays in use: six with FP32/FP64 fields, two with 16-bit
int state_rings(const virs_ctx *c) { return is16(c) ? 2 : 6; }
// bytes of all filter-state rings (one snapshot's share)
size_t state_total(const virs_ctx *c)
{
    size_t b = 0;
    for (int k = 0; k < state_rings(c); ++k) b += state_bytes(c, k);
    return b;
}

size_t elem_size(int prec)
{
    switch (prec) {
    case VIRS_FP16:
    case VIRS_BF16: return 2;
    case VIRS_FP64: return 8;
    default: return 4;
    }
}

int zero_fields(virs_ctx *c, void *const *f)
{
    for (int k = 0; k < 3; ++k)
        if (f[k]) HIPC(c, hipMemsetAsync(f[k], 0, c->esz * (size_t)c->N, c->stream));
    return VIRS_OK;
}

int reset_run(virs_ctx *c)
{
    int rc = zero_fields(c, c->f);
    if (rc) return rc;
    for (int k = 0; k < state_rings(c); ++k)
        if (c->st[k] && c->nb) HIPC(c, hipMemsetAsync(c->st[k], 0, state_bytes(c, k), c->stream));
    hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(64), 0, c->stream, c->ctl, 0, 0LL, 1);
    HIPC(c, hipGetLastError());
    c->ip = 0;
    c->iq = 1;
    c->ir = 2;
    rc = stash_boundary(c, 0);
    if (rc) return rc;
    c->steps = c->counted = 0;
    c->stopped = false;
    c->percent = 0.f;
    c->restart_needed = false;
    drop_graphs(c);
    return VIRS_OK;
}

// Stepping and field reads after virs_adjoint, whose recomputation leaves
// the fields mid-run, need a reset first.
int check_not_mid_adjoint(virs_ctx *c)
{
    if (c->restart_needed)
        return fail(c, VIRS_ERR_STATE,
                    "virs_adjoint left the fields mid-run: call virs_start_simulation (or "
                    "virs_forward_record) before stepping or reading fields");
    return VIRS_OK;
}

} // namespace

extern "C" {

int virs_desc_default(virs_desc *d)
{
    if (!d) return VIRS_ERR_ARG;
    std::memset(d, 0, sizeof *d);
    d->h = 343.f / (22050.f * 2.f * (float)2);
    d->c = 343.f;
    d->dt = 0.f;
    d->precision = VIRS_FP32;
    d->sponge_width = 0;
    d->sponge_sigma = 0.f;
    d->device = 0;
    d->use_graph = 0;
    d->dense_blocks = 0;
    d->fuse_boundary = 0;
    d->rounding = VIRS_ROUND_STOCHASTIC;
    return VIRS_OK;
}

int virs_create(const virs_desc *d, virs_ctx **out)
{
    if (!d || !out) return VIRS_ERR_ARG;
    *out = nullptr;
    auto *c = new virs_ctx();
    c->d = *d;
    c->nx = d->nx;
    c->ny = d->ny;
    c->nz = d->nz;
    if (c->nx < 3 || c->ny < 3 || c->nz < 3 || c->nx > (1 << 30) || c->ny > (1 << 30) ||
        c->nz > (1 << 30) || !(d->h > 0.f) || !(d->c > 0.f) || d->precision < 0 ||
        d->precision > 3 || d->sponge_width < 0 || d->sponge_width > VIRS_MAX_SPONGE ||
        (unsigned long long)c->nx * (unsigned long long)c->ny > VIRS_MAX_CELLS / (unsigned long long)c->nz) {
        delete c;
        return VIRS_ERR_ARG;
    }
    c->N = c->nx * c->ny * c->nz;
    c->esz = elem_size(d->precision);
    int maxv = d->precision == VIRS_FP64 ? 2 : d->precision == VIRS_FP32 ? 4 : 8;
    if (d->vector_width > 0) {
        if ((d->vector_width & (d->vector_width - 1)) || d->vector_width > maxv || c->nx % d->vector_width) {
            delete c;
            return VIRS_ERR_ARG;
        }
        maxv = d->vector_width;
    }
    c->vw = 1;
    for (int v = maxv; v > 1; v >>= 1)
        if (c->nx % v == 0) {
            c->vw = v;
            break;
        }
    // Constants exactly as simulator_gpu.cu:195-198 and :147 compute them.
    const float cs = d->c, h = d->h;
    c->dt = d->dt > 0.f ? d->dt : 0.5f * h / (cs * std::sqrt(3.f));
    c->c2dt2 = cs * cs * c->dt * c->dt;
    c->ih2 = 1.f / (h * h);
    c->kconst = (cs * c->dt) / h * 0.95f;
    // Sponge profile sigma(d) = sigma_max ((W-d)/W)^2 with the centred
    // damping update R = (2P - (1-s)Q + L)/(1+s), s = sigma/2, which stays
    // stable up to the Courant limit.  Outside the layer a,b,g = 2,1,1.
    c->sp.w = d->sponge_width;
    for (int k = 0; k < d->sponge_width; ++k) {
        const double t = (double)(d->sponge_width - k) / (double)d->sponge_width;
        const double s = 0.5 * (double)d->sponge_sigma * t * t;
        c->sp.a[k] = (float)(2.0 / (1.0 + s));
        c->sp.b[k] = (float)((1.0 - s) / (1.0 + s));
        c->sp.g[k] = (float)(1.0 / (1.0 + s));
    }
    c->nchunks_plane = (int)((c->nx * c->ny / c->vw + kBlock - 1) / kBlock);
    hipError_t e = hipSetDevice(d->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, d->device);
    if (e != hipSuccess) {
        delete c;
        return VIRS_ERR_HIP;
    }
    int rc = static_scratch(c, d->device, &c->scratch_dev);
    if (!rc && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = VIRS_ERR_HIP;
    for (int k = 0; k < 3 && !rc; ++k) rc = dalloc(c, &c->f[k], c->esz * (size_t)c->N);
    if (!rc) rc = dalloc(c, &c->code, code_bytes(c));
    if (!rc) rc = dalloc(c, &c->ctl, sizeof(StepCtl));
    if (!rc) rc = dalloc(c, &c->ctl_adj, sizeof(StepCtl));
    if (!rc) rc = dalloc(c, &c->soft_idx, sizeof(long long) * kMaxPoints);
    if (!rc) rc = dalloc(c, &c->soft_bpos, sizeof(long long) * kMaxPoints);
    if (!rc) rc = dalloc(c, &c->rcv_idx, sizeof(long long) * kMaxPoints);
    if (!rc) rc = dalloc(c, &c->rcv_bpos, sizeof(long long) * kMaxPoints);
    if (!rc) rc = dalloc(c, &c->rcv_aux, sizeof(int) * 4 * kMaxPoints);
    if (!rc) rc = dalloc(c, &c->neg, sizeof(long long) * kMaxPoints);
    if (!rc && hipMemset(c->neg, 0xff, sizeof(long long) * kMaxPoints) != hipSuccess) rc = VIRS_ERR_HIP;
    if (!rc) rc = reset_run(c);
    if (!rc) {
        hipLaunchKernelGGL(k_ctl_es, dim3(1), dim3(64), 0, c->stream, c->ctl, 0, c->es_target);
        hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(64), 0, c->stream, c->ctl_adj, 0, 0LL, 1);
        hipLaunchKernelGGL(k_ctl_es, dim3(1), dim3(64), 0, c->stream, c->ctl_adj, 0, c->es_target);
        if (hipGetLastError() != hipSuccess) rc = VIRS_ERR_HIP;
    }
    const int zero_code = is16(c) ? (int)(kNib16Zero | kNib16Zero << 4) : (int)(kTypeZero << 6);
    if (!rc && hipMemsetAsync(c->code, zero_code, code_bytes(c), c->stream) != hipSuccess) rc = VIRS_ERR_HIP;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = VIRS_ERR_HIP;
    if (rc) {
        virs_destroy(c);
        return rc;
    }
    *out = c;
    return VIRS_OK;
}

int virs_destroy(virs_ctx *c)
{
    if (!c) return VIRS_ERR_ARG;
    (void)hipSetDevice(c->d.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->snap_stream) (void)hipStreamSynchronize(c->snap_stream);
    drop_graphs(c);
    for (auto &p : c->f) dfree(p);
    for (auto &p : c->af) dfree(p);
    for (auto &p : c->st) dfree(p);
    for (auto &p : c->hv) dfree(p);
    for (auto &p : c->ast) dfree(p);
    for (auto &p : c->cp_field) dfree(p);
    for (auto &p : c->cp_state) dfree(p);
    dfree(c->code);
    dfree(c->normals);
    dfree(c->bidx);
    dfree(c->bmat);
    dfree(c->coeffs);
    dfree(c->ctl);
    dfree(c->ctl_adj);
    dfree(c->hard_sig);
    dfree(c->soft_idx);
    dfree(c->soft_bpos);
    dfree(c->soft_sig);
    dfree(c->rcv_idx);
    dfree(c->rcv_bpos);
    dfree(c->rcv_aux);
    dfree(c->neg);
    dfree(c->wtab);
    dfree(c->ir_buf);
    for (int k = 0; k < 2; ++k) {
        dfree(c->trace_d[k]);
        dfree(c->trace_y[k]);
        if (c->ev_rec[k]) (void)hipEventDestroy(c->ev_rec[k]);
        if (c->ev_rev[k]) (void)hipEventDestroy(c->ev_rev[k]);
    }
    dfree(c->gsig);
    dfree(c->grad_part);
    dfree(c->snap_dev);
    if (c->snap_host) (void)hipHostFree(c->snap_host);
    for (auto &e : c->snap_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->snap_gather)
        if (e) (void)hipEventDestroy(e);
    if (c->snap_stream) (void)hipStreamDestroy(c->snap_stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VIRS_OK;
}

const char *virs_last_error(const virs_ctx *c) { return c ? c->err.c_str() : "null context"; }

int virs_get_constants(const virs_ctx *c, float *dt, float *c2dt2, float *ih2, float *kconst)
{
    if (!c) return VIRS_ERR_ARG;
    if (dt) *dt = c->dt;
    if (c2dt2) *c2dt2 = c->c2dt2;
    if (ih2) *ih2 = c->ih2;
    if (kconst) *kconst = c->kconst;
    return VIRS_OK;
}

// Soft sources must sit on cells that update (air, boundary shell or a hard
// cell), not on walls or the domain border: those hold zero in 16-bit
// storage and are never read by FP32/FP64 neighbours.  Reads the cell codes
// of the (at most kMaxPoints) source cells back from the device.
int check_soft_cells(virs_ctx *c, const long long *idx, int nsrc)
{
    if (!c->have_flags) return VIRS_OK;
    const bool s16 = c->d.precision == VIRS_FP16 || c->d.precision == VIRS_BF16;
    for (int k = 0; k < nsrc; ++k) {
        const long long i = idx[k];
        uint8_t b = 0;
        HIPC(c, hipMemcpy(&b, c->code + (s16 ? i >> 1 : i), 1, hipMemcpyDeviceToHost));
        const bool zero = s16 ? ((b >> (4 * (i & 1))) & 0xFu) == kNib16Zero : ((b >> 6) & 3u) == 1u;
        if (zero) return fail(c, VIRS_ERR_ARG, "source %d (cell %lld) sits on a wall or border cell", k, i);
    }
    return VIRS_OK;
}

int virs_upload_flags(virs_ctx *c, const uint8_t *flags)
{
    if (!c || !flags) return VIRS_ERR_ARG;
    for (long long i = 0; i < c->N; ++i)
        if (flags[i] > 3) return fail(c, VIRS_ERR_ARG, "flag value %d at %lld (expected 0..3)", flags[i], i);
    HIPC(c, hipSetDevice(c->d.device));
    uint8_t *dflags = nullptr;
    int rc = dalloc(c, &dflags, (size_t)c->N);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(dflags, flags, (size_t)c->N, hipMemcpyHostToDevice, c->stream);
    const bool s16 = c->d.precision == VIRS_FP16 || c->d.precision == VIRS_BF16;
    if (e == hipSuccess) {
        if (s16)
            hipLaunchKernelGGL(k_build_code16, dim3(4096), dim3(256), 0, c->stream, dflags, c->code,
                               (int)c->nx, (int)c->ny, (int)c->nz);
        else
            hipLaunchKernelGGL(k_build_code, dim3(4096), dim3(256), 0, c->stream, dflags, c->code,
                               (int)c->nx, (int)c->ny, (int)c->nz);
        e = hipGetLastError();
    }
    // 16-bit storage keeps walls and border cells at zero in every field
    for (int k = 0; k < 3 && s16 && e == hipSuccess; ++k) {
        if (c->d.precision == VIRS_FP16)
            hipLaunchKernelGGL(k_zero_walls<f16>, dim3(4096), dim3(256), 0, c->stream, (f16 *)c->f[k], c->code, c->N);
        else
            hipLaunchKernelGGL(k_zero_walls<bf16>, dim3(4096), dim3(256), 0, c->stream, (bf16 *)c->f[k], c->code, c->N);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dflags);
    if (e != hipSuccess) return fail(c, VIRS_ERR_HIP, "flag upload: %s", hipGetErrorString(e));
    c->have_flags = true;
    c->h_flag2.clear();
    c->n_hard = 0;
    const long long nx = c->nx, ny = c->ny, nz = c->nz;
    for (long long z = 1; z < nz - 1; ++z)
        for (long long y = 1; y < ny - 1; ++y)
            for (long long x = 1; x < nx - 1; ++x) {
                const long long i = (z * ny + y) * nx + x;
                if (flags[i] == 2) c->h_flag2.push_back((uint32_t)i);
                c->n_hard += flags[i] == 3;
            }
    rc = check_soft_cells(c, c->h_soft_idx.data(), c->nsoft);
    if (rc) return rc;
    c->rec_steps = 0; // a recorded run used the old cell codes
    drop_graphs(c);
    return refresh_fusion(c);
}

int virs_upload_normals(virs_ctx *c, const uint8_t *normals)
{
    if (!c || !normals) return VIRS_ERR_ARG;
    HIPC(c, hipSetDevice(c->d.device));
    if (!c->normals) {
        int rc = dalloc(c, &c->normals, (size_t)c->N);
        if (rc) return rc;
    }
    HIPC(c, hipMemcpy(c->normals, normals, (size_t)c->N, hipMemcpyHostToDevice));
    return VIRS_OK;
}

int virs_upload_boundary(virs_ctx *c, const uint32_t *idx, const uint8_t *mat, int64_t nb)
{
    if (!c || nb < 0 || (nb > 0 && !idx)) return VIRS_ERR_ARG;
    for (int64_t e = 0; e < nb; ++e) {
        if ((long long)idx[e] >= c->N) return fail(c, VIRS_ERR_ARG, "boundary index %u out of range", idx[e]);
        if (e && idx[e] <= idx[e - 1]) return fail(c, VIRS_ERR_ARG, "boundary list not strictly ascending at %lld", (long long)e);
        if (mat && c->nmat && mat[e] >= c->nmat) return fail(c, VIRS_ERR_ARG, "material %d >= %d", mat[e], c->nmat);
    }
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    dfree(c->bidx);
    dfree(c->bmat);
    for (auto &p : c->st) dfree(p);
    for (auto &p : c->hv) dfree(p);
    // the adjoint's list-sized buffers (checkpointed filter states, traces,
    // reverse filter states) follow the new list length at the next record
    // and adjoint; a recorded run no longer matches the list
    for (auto &p : c->cp_state) dfree(p);
    c->cp_state.clear();
    for (int k = 0; k < 2; ++k) {
        dfree(c->trace_d[k]);
        dfree(c->trace_y[k]);
    }
    c->trace_slots = 0;
    for (auto &p : c->ast) dfree(p);
    c->rec_steps = 0;
    c->nb = nb;
    c->h_bidx.assign(idx, idx + nb);
    if (mat) c->h_bmat.assign(mat, mat + nb);
    else c->h_bmat.clear();
    int rc = dalloc(c, &c->bidx, sizeof(uint32_t) * nb);
    for (int k = 0; k < state_rings(c) && !rc; ++k) rc = dalloc(c, &c->st[k], state_bytes(c, k));
    for (int k = 0; k < 2 && !rc; ++k) rc = dalloc(c, &c->hv[k], c->esz * (size_t)std::max<long long>(1, nb));
    // the material array always exists on the device (zeros without one):
    // the list launch reads it unconditionally
    if (!rc) rc = dalloc(c, &c->bmat, (size_t)std::max<long long>(1, nb));
    if (rc) return rc;
    if (nb) HIPC(c, hipMemcpy(c->bidx, idx, sizeof(uint32_t) * nb, hipMemcpyHostToDevice));
    if (nb && mat) HIPC(c, hipMemcpy(c->bmat, mat, (size_t)nb, hipMemcpyHostToDevice));
    if (nb && !mat) HIPC(c, hipMemset(c->bmat, 0, (size_t)nb));
    for (int k = 0; k < state_rings(c); ++k)
        if (nb) HIPC(c, hipMemset(c->st[k], 0, state_bytes(c, k)));
    rc = stash_boundary(c, c->steps);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    rc = refresh_fusion(c);
    if (rc) return rc;
    return upload_points(c);
}

int virs_set_materials(virs_ctx *c, const double *coeffs, int32_t nmat)
{
    if (!c || !coeffs || nmat <= 0 || nmat > 255) return VIRS_ERR_ARG;
    for (size_t e = 0; e < c->h_bmat.size(); ++e)
        if (c->h_bmat[e] >= nmat) return fail(c, VIRS_ERR_ARG, "boundary material %d >= %d", c->h_bmat[e], nmat);
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    dfree(c->coeffs);
    int rc = dalloc(c, &c->coeffs, sizeof(double) * 5 * nmat);
    if (rc) return rc;
    HIPC(c, hipMemcpy(c->coeffs, coeffs, sizeof(double) * 5 * nmat, hipMemcpyHostToDevice));
    c->nmat = nmat;
    c->rec_steps = 0; // a recorded run used the old coefficients
    drop_graphs(c);
    return VIRS_OK;
}

int virs_set_hard_source(virs_ctx *c, const double *sig, int64_t len)
{
    if (!c || len < 0 || (sig && len == 0)) return VIRS_ERR_ARG;
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    dfree(c->hard_sig);
    c->hard_ref = sig ? 0 : 1;
    c->hard_len = sig ? len : 0;
    if (sig) {
        int rc = dalloc(c, &c->hard_sig, sizeof(double) * len);
        if (rc) return rc;
        HIPC(c, hipMemcpy(c->hard_sig, sig, sizeof(double) * len, hipMemcpyHostToDevice));
    }
    c->rec_steps = 0; // a recorded run used the old source
    drop_graphs(c);
    return VIRS_OK;
}

int virs_set_soft_sources(virs_ctx *c, const int64_t *idx, const double *sig, int32_t nsrc, int64_t len)
{
    if (!c || nsrc < 0 || nsrc > kMaxPoints || len < 0 || (nsrc && (!idx || !sig || len == 0)))
        return VIRS_ERR_ARG;
    for (int k = 0; k < nsrc; ++k) {
        if (idx[k] < 0 || idx[k] >= c->N) return fail(c, VIRS_ERR_ARG, "source %d index out of range", k);
        for (int j = 0; j < k; ++j)
            if (idx[j] == idx[k]) return fail(c, VIRS_ERR_ARG, "sources %d and %d share a cell", j, k);
    }
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    {
        std::vector<long long> cells(idx, idx + nsrc);
        int rc = check_soft_cells(c, cells.data(), nsrc);
        if (rc) return rc;
    }
    dfree(c->soft_sig);
    c->nsoft = nsrc;
    c->soft_len = nsrc ? len : 0;
    c->h_soft_idx.assign(idx, idx + nsrc);
    c->h_soft_sig.assign(sig, sig + (size_t)nsrc * (size_t)c->soft_len);
    if (nsrc) {
        int rc = dalloc(c, &c->soft_sig, sizeof(double) * nsrc * len);
        if (rc) return rc;
        HIPC(c, hipMemcpy(c->soft_sig, sig, sizeof(double) * nsrc * len, hipMemcpyHostToDevice));
    }
    c->rec_steps = 0; // a recorded run used the old sources
    drop_graphs(c);
    return upload_points(c);
}

int virs_set_receivers(virs_ctx *c, const int64_t *idx, int32_t nrcv)
{
    if (!c || nrcv < 0 || nrcv > kMaxPoints || (nrcv && !idx)) return VIRS_ERR_ARG;
    for (int k = 0; k < nrcv; ++k)
        if (idx[k] < 0 || idx[k] >= c->N) return fail(c, VIRS_ERR_ARG, "receiver %d index out of range", k);
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    c->nrcv = nrcv;
    c->h_rcv_idx.assign(idx, idx + nrcv);
    dfree(c->ir_buf);
    c->ir_cap = 0;
    drop_graphs(c);
    return upload_points(c);
}

int virs_set_early_stop(virs_ctx *c, int32_t enable, float target)
{
    if (!c) return VIRS_ERR_ARG;
    c->es_enable = enable;
    c->es_target = target > 0.f ? target : 0.0009f;
    HIPC(c, hipSetDevice(c->d.device));
    hipLaunchKernelGGL(k_ctl_es, dim3(1), dim3(64), 0, c->stream, c->ctl, c->es_enable, c->es_target);
    HIPC(c, hipGetLastError());
    return VIRS_OK;
}

// Switches the device-side RMS rule off (timing, recording, adjoint) or
// back to the context's setting.
static int device_early_stop(virs_ctx *c, bool on)
{
    hipLaunchKernelGGL(k_ctl_es, dim3(1), dim3(64), 0, c->stream, c->ctl, on ? c->es_enable : 0, c->es_target);
    HIPC(c, hipGetLastError());
    return VIRS_OK;
}

// Holds the RMS stop off while every step must run (timing, recording, the
// adjoint's recomputation).  resume() restores the context's setting; an
// error return before it restores the setting as well, after the work queued
// on `other` (a second stream whose steps read the setting) has finished.
struct EarlyStopPause {
    virs_ctx *c;
    hipStream_t other;
    bool paused = false;
    explicit EarlyStopPause(virs_ctx *ctx, hipStream_t o = nullptr) : c(ctx), other(o) {}
    int pause()
    {
        const int rc = device_early_stop(c, false);
        paused = rc == VIRS_OK;
        return rc;
    }
    int resume()
    {
        paused = false;
        return device_early_stop(c, true);
    }
    ~EarlyStopPause()
    {
        if (!paused) return;
        if (other) (void)hipStreamSynchronize(other);
        (void)device_early_stop(c, true);
    }
    EarlyStopPause(const EarlyStopPause &) = delete;
    EarlyStopPause &operator=(const EarlyStopPause &) = delete;
};

int virs_start_simulation(virs_ctx *c, int64_t nsteps)
{
    if (!c || nsteps < 0) return VIRS_ERR_ARG;
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    dfree(c->ir_buf);
    c->ir_cap = nsteps;
    const size_t bytes = sizeof(double) * (size_t)std::max(1, c->nrcv) * (size_t)std::max<long long>(1, nsteps);
    int rc = dalloc(c, &c->ir_buf, bytes);
    if (rc) return rc;
    HIPC(c, hipMemsetAsync(c->ir_buf, 0, bytes, c->stream));
    rc = reset_run(c);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return VIRS_OK;
}

// Captures `steps` plain steps from the current rotation phase into *out.
static int build_graph(virs_ctx *c, int steps, hipGraphExec_t *out)
{
    const int sp = c->ip, sq = c->iq, sr = c->ir;
    hipGraph_t g = nullptr;
    HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = VIRS_OK;
    for (int s = 0; s < steps && !rc; ++s) rc = launch_step(c);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->ip = sp;
    c->iq = sq;
    c->ir = sr;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(c, VIRS_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(c, VIRS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    return VIRS_OK;
}

// How run_steps splits n steps when the context uses graphs: `full`
// replays of the kGraphSteps-step graph, then one graph of `tail` steps.
static void graph_plan(long long n, long long &full, int &tail)
{
    full = n / kGraphSteps;
    tail = (int)(n % kGraphSteps);
}

// Makes sure the graphs that run_steps(n) replays from the current rotation
// phase exist (whole replays keep the phase, so the tail starts from it too).
static int ensure_graphs_for(virs_ctx *c, long long n)
{
    long long full;
    int tail;
    graph_plan(n, full, tail);
    const int ph = c->ip;
    if (full > 0 && !c->graph[ph]) {
        const int rc = build_graph(c, kGraphSteps, &c->graph[ph]);
        if (rc) return rc;
    }
    if (tail > 0 && (!c->tail_graph[ph] || c->tail_steps[ph] != tail)) {
        if (c->tail_graph[ph]) (void)hipGraphExecDestroy(c->tail_graph[ph]);
        c->tail_graph[ph] = nullptr;
        c->tail_steps[ph] = 0;
        const int rc = build_graph(c, tail, &c->tail_graph[ph]);
        if (rc) return rc;
        c->tail_steps[ph] = tail;
    }
    return VIRS_OK;
}

// Applies the buffer rotation of k steps to the host's indices.
static void rotate(virs_ctx *c, long long k)
{
    for (long long j = 0; j < k % 3; ++j) {
        const int ip = c->ip;
        c->ip = c->ir;
        c->ir = c->iq;
        c->iq = ip;
    }
}

// n plain steps (receivers on, no trace) on the context's stream.  With
// graphs: replays of the kGraphSteps-step graph of the current rotation
// phase, then one graph of the remaining steps; otherwise single launches.
static int run_steps(virs_ctx *c, long long n)
{
    if (c->d.use_graph && n > 0) {
        int rc = ensure_graphs_for(c, n);
        if (rc) return rc;
        long long full;
        int tail;
        graph_plan(n, full, tail);
        for (long long k = 0; k < full; ++k) {
            HIPC(c, hipGraphLaunch(c->graph[c->ip], c->stream));
            c->steps += kGraphSteps;
        }
        if (tail > 0) {
            HIPC(c, hipGraphLaunch(c->tail_graph[c->ip], c->stream));
            c->steps += tail;
            rotate(c, tail);
        }
        return VIRS_OK;
    }
    for (; n > 0; --n) {
        const int rc = launch_step(c);
        if (rc) return rc;
        ++c->steps;
    }
    return VIRS_OK;
}

int virs_step(virs_ctx *c, int64_t nsteps, int32_t *keep_going)
{
    int rc = check_ready(c);
    if (!rc) rc = check_not_mid_adjoint(c);
    if (rc) return rc;
    if (nsteps < 0) return VIRS_ERR_ARG;
    if (keep_going) *keep_going = c->stopped ? 0 : 1;
    if (c->stopped) return VIRS_OK;
    rc = run_steps(c, nsteps);
    if (rc) return rc;
    c->counted = c->steps;
    if (c->es_enable && c->nrcv > 0 && c->steps > (long long)kRmsMinStep + 1) {
        // The device evaluates the RMS rule at the end of every step and
        // skips every step after the one where it fires; read its verdict.
        // No step up to kRmsMinStep can fire, so the stream is waited on
        // only from then on.
        long long stop_at = 0;
        float percent = 0.f;
        HIPC(c, hipMemcpyAsync(&stop_at, &c->ctl->stop_at, sizeof stop_at, hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipMemcpyAsync(&percent, &c->ctl->percent, sizeof percent, hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        c->percent = percent;
        if (stop_at != 0x7fffffffffffffffll) {
            // the steps launched after stop_at left the fields untouched:
            // rewind the rotation to just after stop_at and point its slot
            // of the step counter back at stop_at + 1, clearing the stop bit
            const long long extra = c->steps - (stop_at + 1);
            for (long long k = 0; k < extra % 3; ++k) {
                const int ip = c->ip;
                c->ip = c->iq;
                c->iq = c->ir;
                c->ir = ip;
            }
            c->steps = c->counted = stop_at + 1;
            hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(64), 0, c->stream, c->ctl, c->ip, c->steps, 0);
            HIPC(c, hipGetLastError());
            HIPC(c, hipStreamSynchronize(c->stream));
            c->stopped = true;
            if (keep_going) *keep_going = 0;
        }
    }
    return VIRS_OK;
}

int64_t virs_steps_done(const virs_ctx *c) { return c ? c->counted : -1; }

float virs_percent_to_target_rms(const virs_ctx *c) { return c ? c->percent : 0.f; }

int virs_synchronize(virs_ctx *c)
{
    if (!c) return VIRS_ERR_ARG;
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    int err = 0;
    HIPC(c, hipMemcpy(&err, &c->ctl->err, sizeof(int), hipMemcpyDeviceToHost));
    if (err)
        return fail(c, VIRS_ERR_STATE,
                    "the run diverged: a receiver sample, a corrected boundary value or the 16-bit "
                    "scale estimate was not finite");
    return VIRS_OK;
}

int virs_read_receivers_f64(virs_ctx *c, double *out)
{
    if (!c || (!out && c->nrcv)) return VIRS_ERR_ARG;
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    const long long k = std::min(c->counted, c->ir_cap);
    for (int r = 0; r < c->nrcv; ++r)
        if (k) HIPC(c, hipMemcpy(out + (long long)r * k, c->ir_buf + (long long)r * c->ir_cap,
                                 sizeof(double) * k, hipMemcpyDeviceToHost));
    return VIRS_OK;
}

int virs_read_receivers(virs_ctx *c, float *out)
{
    if (!c) return VIRS_ERR_ARG;
    const long long k = std::min(c->counted, c->ir_cap);
    std::vector<double> tmp((size_t)std::max<long long>(1, k * c->nrcv));
    int rc = virs_read_receivers_f64(c, tmp.data());
    if (rc) return rc;
    for (long long i = 0; i < k * c->nrcv; ++i) out[i] = (float)tmp[i];
    return VIRS_OK;
}

int virs_field_exponent(virs_ctx *c, int32_t which, int32_t *e)
{
    if (!c || !e || which < 0 || which > 2) return VIRS_ERR_ARG;
    *e = 0;
    if (int rc = check_not_mid_adjoint(c)) return rc;
    if (c->d.precision != VIRS_FP16 && c->d.precision != VIRS_BF16) return VIRS_OK;
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    StepCtl h;
    HIPC(c, hipMemcpy(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost));
    *e = h.exp[(h.nring[c->ip] - 1 - which) & 3];
    return VIRS_OK;
}

static const void *field_buffer(const virs_ctx *c, int which)
{
    return c->f[which == 0 ? c->ip : which == 1 ? c->iq : c->ir];
}

int virs_read_field_f64(virs_ctx *c, int32_t which, double *out)
{
    if (!c || !out || which < 0 || which > 2) return VIRS_ERR_ARG;
    int e = 0;
    int rc = virs_field_exponent(c, which, &e);
    if (rc) return rc;
    double *tmp = nullptr;
    rc = dalloc(c, &tmp, sizeof(double) * c->N);
    if (rc) return rc;
    const void *src = field_buffer(c, which);
    const int g = 4096;
    switch (c->d.precision) {
    case VIRS_FP16: hipLaunchKernelGGL(k_to_f64<f16>, dim3(g), dim3(256), 0, c->stream, (const f16 *)src, tmp, c->N, e); break;
    case VIRS_BF16: hipLaunchKernelGGL(k_to_f64<bf16>, dim3(g), dim3(256), 0, c->stream, (const bf16 *)src, tmp, c->N, e); break;
    case VIRS_FP64: hipLaunchKernelGGL(k_to_f64<double>, dim3(g), dim3(256), 0, c->stream, (const double *)src, tmp, c->N, e); break;
    default: hipLaunchKernelGGL(k_to_f64<float>, dim3(g), dim3(256), 0, c->stream, (const float *)src, tmp, c->N, e); break;
    }
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpyAsync(out, tmp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (err != hipSuccess) return fail(c, VIRS_ERR_HIP, "read field: %s", hipGetErrorString(err));
    return VIRS_OK;
}

int virs_read_field(virs_ctx *c, int32_t which, float *out)
{
    if (!c || !out || which < 0 || which > 2) return VIRS_ERR_ARG;
    if (int rc = check_not_mid_adjoint(c)) return rc;
    if (c->d.precision == VIRS_FP32) {
        HIPC(c, hipSetDevice(c->d.device));
        HIPC(c, hipMemcpyAsync(out, field_buffer(c, which), sizeof(float) * c->N, hipMemcpyDeviceToHost,
                               c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        return VIRS_OK;
    }
    std::vector<double> tmp((size_t)c->N);
    int rc = virs_read_field_f64(c, which, tmp.data());
    if (rc) return rc;
    for (long long i = 0; i < c->N; ++i) out[i] = (float)tmp[i];
    return VIRS_OK;
}

int virs_write_field_f64(virs_ctx *c, int32_t which, const double *in)
{
    if (!c || !in || which < 0 || which > 1) return VIRS_ERR_ARG;
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipStreamSynchronize(c->stream));
    int e = 0;
    if (c->d.precision == VIRS_FP16 || c->d.precision == VIRS_BF16) {
        // pick the exponent for the new field and record it where the step
        // reads it: exp[(n-1)&3] for p^n, exp[(n-2)&3] for p^{n-1}
        double mx = 0.0;
        for (long long i = 0; i < c->N; ++i) mx = std::max(mx, std::fabs(in[i]));
        if (mx > 0.0) {
            int k2;
            std::frexp(mx * 16.0, &k2);
            e = std::min(std::max(14 - k2, -60), 60);
        }
        StepCtl h;
        HIPC(c, hipMemcpy(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost));
        const int slot = (int)((h.nring[c->ip] - 1 - which) & 3);
        h.exp[slot] = e;
        const float fm = (float)mx;
        unsigned bits;
        std::memcpy(&bits, &fm, 4);
        for (int s = 0; s < kSlots; ++s) h.maxabs[slot][s] = bits;
        HIPC(c, hipMemcpy(c->ctl, &h, sizeof h, hipMemcpyHostToDevice));
    }
    double *tmp = nullptr;
    int rc = dalloc(c, &tmp, sizeof(double) * c->N);
    if (rc) return rc;
    void *dst = c->f[which == 0 ? c->ip : c->iq];
    hipError_t err = hipMemcpy(tmp, in, sizeof(double) * c->N, hipMemcpyHostToDevice);
    const int g = 4096;
    if (err == hipSuccess) {
        switch (c->d.precision) {
        case VIRS_FP16:
            hipLaunchKernelGGL(k_from_f64<f16>, dim3(g), dim3(256), 0, c->stream, tmp, (f16 *)dst, c->N, e);
            if (c->have_flags)
                hipLaunchKernelGGL(k_zero_walls<f16>, dim3(g), dim3(256), 0, c->stream, (f16 *)dst, c->code, c->N);
            break;
        case VIRS_BF16:
            hipLaunchKernelGGL(k_from_f64<bf16>, dim3(g), dim3(256), 0, c->stream, tmp, (bf16 *)dst, c->N, e);
            if (c->have_flags)
                hipLaunchKernelGGL(k_zero_walls<bf16>, dim3(g), dim3(256), 0, c->stream, (bf16 *)dst, c->code, c->N);
            break;
        case VIRS_FP64: hipLaunchKernelGGL(k_from_f64<double>, dim3(g), dim3(256), 0, c->stream, tmp, (double *)dst, c->N, e); break;
        default: hipLaunchKernelGGL(k_from_f64<float>, dim3(g), dim3(256), 0, c->stream, tmp, (float *)dst, c->N, e); break;
        }
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (err != hipSuccess) return fail(c, VIRS_ERR_HIP, "write field: %s", hipGetErrorString(err));
    rc = stash_boundary(c, c->steps);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return VIRS_OK;
}

// ---------------------------------------------------------------------------
// Layer snapshots (layerToImage's input, vectorspace.cpp:77-148): one XZ
// plane of the current field instead of the whole field, gathered on the
// device and copied into pinned host memory behind the steps already queued.
// ---------------------------------------------------------------------------

// Allocates the snapshot ring on first use; the pinned buffer comes last, so
// its presence means everything is in place (a failed call can be retried).
static int ensure_snapshots(virs_ctx *c)
{
    if (c->snap_host) return VIRS_OK;
    for (auto &e : c->snap_ev)
        if (!e) HIPC(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto &e : c->snap_gather)
        if (!e) HIPC(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (!c->snap_stream) HIPC(c, hipStreamCreateWithFlags(&c->snap_stream, hipStreamNonBlocking));
    const size_t bytes = (size_t)kSnapSlots * c->nx * c->nz * sizeof(float);
    if (!c->snap_dev) {
        int rc = dalloc(c, &c->snap_dev, bytes);
        if (rc) return rc;
    }
    if (hipHostMalloc((void **)&c->snap_host, bytes, hipHostMallocDefault) != hipSuccess) {
        c->snap_host = nullptr;
        return fail(c, VIRS_ERR_HIP, "pinned snapshot buffer of %zu bytes", bytes);
    }
    return VIRS_OK;
}

int virs_plane_request(virs_ctx *c, int64_t layer, int32_t *ticket)
{
    if (!c || !ticket) return VIRS_ERR_ARG;
    if (layer < 0 || layer >= c->ny)
        return fail(c, VIRS_ERR_ARG, "layer %lld out of range (Ny = %lld)", (long long)layer, c->ny);
    if (int rc = check_not_mid_adjoint(c)) return rc;
    HIPC(c, hipSetDevice(c->d.device));
    int rc = ensure_snapshots(c);
    if (rc) return rc;
    const int slot = c->snap_next % kSnapSlots;
    if (c->snap_ticket[slot] >= 0)
        return fail(c, VIRS_ERR_ARG, "%d snapshots outstanding: fetch ticket %d first", kSnapSlots,
                    c->snap_ticket[slot]);
    const size_t plane = (size_t)c->nx * c->nz;
    float *dst = c->snap_dev + slot * plane;
    const void *src = c->f[c->ip];
    const dim3 grid((unsigned)((c->nx + 255) / 256), (unsigned)std::min<long long>(c->nz, 65535));
    const int nx = (int)c->nx, ny = (int)c->ny, nz = (int)c->nz, y = (int)layer;
    switch (c->d.precision) {
    case VIRS_FP16:
        hipLaunchKernelGGL(k_plane_y<f16>, grid, dim3(256), 0, c->stream, (const f16 *)src, c->ctl, c->ip, dst, nx,
                           ny, nz, y);
        break;
    case VIRS_BF16:
        hipLaunchKernelGGL(k_plane_y<bf16>, grid, dim3(256), 0, c->stream, (const bf16 *)src, c->ctl, c->ip, dst,
                           nx, ny, nz, y);
        break;
    case VIRS_FP64:
        hipLaunchKernelGGL(k_plane_y<double>, grid, dim3(256), 0, c->stream, (const double *)src, c->ctl, c->ip,
                           dst, nx, ny, nz, y);
        break;
    default:
        hipLaunchKernelGGL(k_plane_y<float>, grid, dim3(256), 0, c->stream, (const float *)src, c->ctl, c->ip, dst,
                           nx, ny, nz, y);
        break;
    }
    HIPC(c, hipGetLastError());
    // the copy waits for the gather only; later steps on the main stream do
    // not touch the gathered plane, so they run while it copies
    HIPC(c, hipEventRecord(c->snap_gather[slot], c->stream));
    HIPC(c, hipStreamWaitEvent(c->snap_stream, c->snap_gather[slot], 0));
    HIPC(c, hipMemcpyAsync(c->snap_host + slot * plane, dst, plane * sizeof(float), hipMemcpyDeviceToHost,
                           c->snap_stream));
    HIPC(c, hipEventRecord(c->snap_ev[slot], c->snap_stream));
    c->snap_ticket[slot] = c->snap_next;
    *ticket = c->snap_next;
    c->snap_next = (c->snap_next + 1) & 0x7fffffff; // 2^31 is a multiple of kSnapSlots
    return VIRS_OK;
}

int virs_plane_fetch(virs_ctx *c, int32_t ticket, float *out)
{
    if (!c || !out || ticket < 0) return VIRS_ERR_ARG;
    const int slot = ticket % kSnapSlots;
    if (!c->snap_host || c->snap_ticket[slot] != ticket)
        return fail(c, VIRS_ERR_ARG, "no outstanding snapshot with ticket %d", ticket);
    HIPC(c, hipSetDevice(c->d.device));
    HIPC(c, hipEventSynchronize(c->snap_ev[slot]));
    const size_t plane = (size_t)c->nx * c->nz;
    std::memcpy(out, c->snap_host + slot * plane, plane * sizeof(float));
    c->snap_ticket[slot] = -1;
    return VIRS_OK;
}

int virs_plane_read_noalloc(virs_ctx *c, int64_t layer, float *out)
{
    if (!c || !out) return VIRS_ERR_ARG;
    if (layer < 0 || layer >= c->ny)
        return fail(c, VIRS_ERR_ARG, "layer %lld out of range (Ny = %lld)", (long long)layer, c->ny);
    if (c->nx * c->nz > kStaticPlaneCells)
        return fail(c, VIRS_ERR_ARG, "plane of %lld cells exceeds the static staging of %lld cells",
                    c->nx * c->nz, kStaticPlaneCells);
    if (int rc = check_not_mid_adjoint(c)) return rc;
    HIPC(c, hipSetDevice(c->d.device));
    float *dst = c->scratch_dev;
    const void *src = c->f[c->ip];
    const dim3 grid((unsigned)((c->nx + 255) / 256), (unsigned)std::min<long long>(c->nz, 65535));
    const int nx = (int)c->nx, ny = (int)c->ny, nz = (int)c->nz, y = (int)layer;
    switch (c->d.precision) {
    case VIRS_FP16:
        hipLaunchKernelGGL(k_plane_y<f16>, grid, dim3(256), 0, c->stream, (const f16 *)src, c->ctl, c->ip, dst, nx,
                           ny, nz, y);
        break;
    case VIRS_BF16:
        hipLaunchKernelGGL(k_plane_y<bf16>, grid, dim3(256), 0, c->stream, (const bf16 *)src, c->ctl, c->ip, dst,
                           nx, ny, nz, y);
        break;
    case VIRS_FP64:
        hipLaunchKernelGGL(k_plane_y<double>, grid, dim3(256), 0, c->stream, (const double *)src, c->ctl, c->ip,
                           dst, nx, ny, nz, y);
        break;
    default:
        hipLaunchKernelGGL(k_plane_y<float>, grid, dim3(256), 0, c->stream, (const float *)src, c->ctl, c->ip, dst,
                           nx, ny, nz, y);
        break;
    }
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(out, dst, sizeof(float) * (size_t)c->nx * (size_t)c->nz, hipMemcpyDeviceToHost,
                           c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return VIRS_OK;
}

int virs_memory_bytes(const virs_ctx *c, int64_t *bytes)
{
    if (!c || !bytes) return VIRS_ERR_ARG;
    long long b = 3 * (long long)c->esz * c->N + (long long)code_bytes(c); // fields + cell codes
    if (c->normals) b += c->N;
    b += c->nb * (4 + 2 * (long long)c->esz + 1) + (long long)state_total(c);
    b += (long long)sizeof(double) * std::max(1, c->nrcv) * std::max<long long>(1, c->ir_cap);
    b += (long long)c->cp_field.size() * (long long)c->esz * c->N;
    b += (long long)c->cp_state.size() * (long long)state_total(c);
    b += (c->trace_d[1] ? 4 : 2) * (long long)compute_bytes(c) * c->trace_slots * c->nb; // d and y per buffer
    for (auto p : c->af)
        if (p) b += (long long)c->esz * c->N;
    if (c->snap_host) b += (long long)kSnapSlots * c->nx * c->nz * (long long)sizeof(float);
    if (c->wtab) b += (long long)sizeof(uint2) * c->nchunks_plane * kWaves * c->nz;
    *bytes = b;
    return VIRS_OK;
}

int virs_time_steps(virs_ctx *c, int64_t nsteps, int64_t nsample, virs_timing *out)
{
    int rc = check_ready(c);
    if (!rc) rc = check_not_mid_adjoint(c);
    if (rc) return rc;
    if (nsteps <= 0 || nsample < 0 || !out) return VIRS_ERR_ARG;
    std::memset(out, 0, sizeof *out);
    EarlyStopPause es(c); // every sampled or timed step does its work
    rc = es.pause();
    if (rc) return rc;
    // The timed region's graphs are built first, for the rotation phase it
    // starts in (after the sampling pass), so that nothing sits between the
    // sampling pass and the timed region: the sampled launches are still
    // running when the start event and the first graph are queued, and the
    // timed region does not include the latency of launching onto an idle
    // GPU, which a run of many steps does not have either.
    if (!rc && c->d.use_graph) {
        rotate(c, nsample);
        rc = ensure_graphs_for(c, nsteps);
        rotate(c, (3 - nsample % 3) % 3);
    }
    // Sampling pass: nsample single steps whose two launches record their
    // own start and stop (hipExtLaunchKernel), outside the timed region.
    std::vector<hipEvent_t> ev((size_t)nsample * 4, nullptr);
    for (auto &e : ev)
        if (!rc && hipEventCreate(&e) != hipSuccess) rc = fail(c, VIRS_ERR_HIP, "event create");
    for (long long k = 0; k < nsample && !rc; ++k) {
        rc = launch_step(c, nullptr, nullptr, 0, true, &ev[(size_t)k * 4]);
        if (!rc) ++c->steps;
    }
    // The timed region runs exactly what virs_step runs.
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (!rc && hipEventCreate(&t0) != hipSuccess) rc = fail(c, VIRS_ERR_HIP, "event create");
    if (!rc && hipEventCreate(&t1) != hipSuccess) rc = fail(c, VIRS_ERR_HIP, "event create");
    if (!rc && hipEventRecord(t0, c->stream) != hipSuccess) rc = fail(c, VIRS_ERR_HIP, "event record");
    if (!rc) rc = run_steps(c, nsteps);
    if (!rc && hipEventRecord(t1, c->stream) != hipSuccess) rc = fail(c, VIRS_ERR_HIP, "event record");
    c->counted = c->steps;
    if (!rc) rc = es.resume();
    if (!rc && hipEventSynchronize(t1) != hipSuccess) rc = fail(c, VIRS_ERR_HIP, "event sync");
    double dsum = 0, bsum = 0;
    for (long long j = 0; j < nsample && !rc; ++j) {
        float a = 0, b = 0;
        if (hipEventElapsedTime(&a, ev[j * 4], ev[j * 4 + 1]) != hipSuccess ||
            hipEventElapsedTime(&b, ev[j * 4 + 2], ev[j * 4 + 3]) != hipSuccess)
            rc = fail(c, VIRS_ERR_HIP, "sampled launch events");
        dsum += a;
        bsum += b;
    }
    float tot = 0;
    if (!rc && hipEventElapsedTime(&tot, t0, t1) != hipSuccess) rc = fail(c, VIRS_ERR_HIP, "timed region events");
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (rc) return rc;
    out->total_ms = tot;
    out->dense_ms = nsample ? dsum / (double)nsample : 0.0;
    out->boundary_ms = nsample ? bsum / (double)nsample : 0.0;
    out->sampled_steps = nsample;
    if (c->d.use_graph) {
        long long full;
        int tail;
        graph_plan(nsteps, full, tail);
        out->graph_launches = full + (tail > 0);
        out->graph_steps = full > 0 ? kGraphSteps : 0;
        out->tail_graph_steps = tail;
    } else {
        out->single_steps = nsteps;
    }
    return VIRS_OK;
}

// ---------------------------------------------------------------------------
// Adjoint with checkpointed recomputation.
//
// Forward: snapshots of (p^n, p^{n-1}, filter state) at the start of every
// segment of `seg` steps; the last segment records its boundary trace
// (d^n, y^n per entry) directly.  Reverse, segment by segment from the end:
// the reverse steps of segment k read the trace of segment k while segment
// k-1 is recomputed from its snapshot on a second stream, recording its trace
// into the other buffer.  Events order the two streams: a recomputation
// starts once the reverse sweep that last read its buffer has finished, and a
// reverse sweep once its segment's recomputation has.  Cost: one forward,
// one recomputation of all but the last segment (overlapped with the reverse
// sweep), one reverse sweep.
// ---------------------------------------------------------------------------

// Runs the context's launches on stream s for the lifetime of the object.
struct OnStream {
    virs_ctx *c;
    hipStream_t saved;
    OnStream(virs_ctx *ctx, hipStream_t s) : c(ctx), saved(ctx->stream) { c->stream = s; }
    ~OnStream() { c->stream = saved; }
};

// Trace buffers of `slots` steps: one when a single segment is recorded
// whole (nothing is recomputed), two when recomputation overlaps the sweep.
static int ensure_trace(virs_ctx *c, long long slots, int nbuf)
{
    if (c->trace_slots >= slots && c->trace_d[0] && (nbuf == 1 || c->trace_d[1])) return VIRS_OK;
    for (int k = 0; k < 2; ++k) {
        dfree(c->trace_d[k]);
        dfree(c->trace_y[k]);
    }
    c->trace_slots = 0;
    const size_t bytes = compute_bytes(c) * slots * std::max<long long>(1, c->nb);
    int rc = VIRS_OK;
    for (int k = 0; k < nbuf && !rc; ++k) {
        rc = dalloc(c, &c->trace_d[k], bytes);
        if (!rc) rc = dalloc(c, &c->trace_y[k], bytes);
    }
    if (rc) return rc;
    c->trace_slots = slots;
    return VIRS_OK;
}

// Trace slots 0 and 1 of buffer buf from the filter state before step s0.
static int seed_trace(virs_ctx *c, int buf, long long s0)
{
    if (!c->nb) return VIRS_OK;
    const int a = (int)(s0 & 1), b = a ^ 1; // ring slots of steps s0-2 and s0-1
    const double *y1 = (const double *)c->st[2 + b], *y2 = (const double *)c->st[2 + a];
    if (c->d.precision == VIRS_FP64)
        hipLaunchKernelGGL(k_trace_seed<double>, dim3(grid_for(c->nb)), dim3(256), 0, c->stream,
                           (const double *)c->st[b], (const double *)c->st[a], y1, y2,
                           (double *)c->trace_d[buf], (double *)c->trace_y[buf], c->nb);
    else
        hipLaunchKernelGGL(k_trace_seed<float>, dim3(grid_for(c->nb)), dim3(256), 0, c->stream,
                           (const float *)c->st[b], (const float *)c->st[a], y1, y2,
                           (float *)c->trace_d[buf], (float *)c->trace_y[buf], c->nb);
    HIPC(c, hipGetLastError());
    return VIRS_OK;
}

static int save_checkpoint(virs_ctx *c, int k)
{
    HIPC(c, hipMemcpyAsync(c->cp_field[2 * k], c->f[c->ip], c->esz * c->N, hipMemcpyDeviceToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->cp_field[2 * k + 1], c->f[c->iq], c->esz * c->N, hipMemcpyDeviceToDevice, c->stream));
    size_t off = 0;
    for (int j = 0; j < state_rings(c) && c->nb; ++j) {
        HIPC(c, hipMemcpyAsync(c->cp_state[k] + off, c->st[j], state_bytes(c, j), hipMemcpyDeviceToDevice,
                               c->stream));
        off += state_bytes(c, j);
    }
    return VIRS_OK;
}

// Recomputes segment k from its snapshot on the current launch stream,
// recording its trace into buffer k & 1.
static int recompute_segment(virs_ctx *c, int k)
{
    const long long s0 = (long long)k * c->seg, s1 = std::min(c->rec_steps, s0 + c->seg);
    const int buf = k & 1;
    c->ip = 0;
    c->iq = 1;
    c->ir = 2;
    HIPC(c, hipMemcpyAsync(c->f[c->ip], c->cp_field[2 * k], c->esz * c->N, hipMemcpyDeviceToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->f[c->iq], c->cp_field[2 * k + 1], c->esz * c->N, hipMemcpyDeviceToDevice, c->stream));
    size_t off = 0;
    for (int j = 0; j < state_rings(c) && c->nb; ++j) {
        HIPC(c, hipMemcpyAsync(c->st[j], c->cp_state[k] + off, state_bytes(c, j), hipMemcpyDeviceToDevice,
                               c->stream));
        off += state_bytes(c, j);
    }
    hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(64), 0, c->stream, c->ctl, 0, s0, 0);
    HIPC(c, hipGetLastError());
    int rc = stash_boundary(c, s0);
    if (!rc) rc = seed_trace(c, buf, s0);
    for (long long s = s0; s < s1 && !rc; ++s) rc = launch_step(c, c->trace_d[buf], c->trace_y[buf], s0, false);
    return rc;
}

// The RMS stop stays off while a run is recorded for the adjoint and while
// the adjoint recomputes it: every recorded step is kept.  on = 1 restores
// the context's setting.
int virs_forward_record(virs_ctx *c, int64_t nsteps, int32_t ncp)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->d.precision != VIRS_FP32 && c->d.precision != VIRS_FP64)
        return fail(c, VIRS_ERR_UNSUPPORTED, "adjoint needs FP32 or FP64 fields");
    if (c->nmat > 8) return fail(c, VIRS_ERR_UNSUPPORTED, "adjoint supports up to 8 materials");
    if (nsteps <= 0 || ncp < 0) return VIRS_ERR_ARG;
    if (c->ir_cap < nsteps) return fail(c, VIRS_ERR_STATE, "receiver buffer shorter than nsteps");
    if (ncp == 0) {
        // automatic: record the whole boundary trace (no recomputation) when
        // it takes at most a quarter of the free device memory, else sqrt(N)
        // checkpoints
        size_t free_b = 0, total_b = 0;
        HIPC(c, hipMemGetInfo(&free_b, &total_b));
        const double whole = 2.0 * (double)compute_bytes(c) * (double)(nsteps + 2) *
                             (double)std::max<long long>(1, c->nb);
        ncp = whole <= 0.25 * (double)free_b ? 1 : (int)std::ceil(std::sqrt((double)nsteps));
    }
    ncp = (int)std::min<long long>(ncp, nsteps);
    const long long seg = (nsteps + ncp - 1) / ncp;
    ncp = (int)((nsteps + seg - 1) / seg);
    c->rec_steps = 0; // a new recording replaces the last one, also when it fails
    // (re)allocate snapshots: the new ones are checked against the free
    // device memory first, so an impossible request fails at once; a failure
    // part-way (memory taken meanwhile) leaves no snapshots behind
    if ((int)c->cp_state.size() != ncp) {
        for (auto &p : c->cp_field) dfree(p);
        for (auto &p : c->cp_state) dfree(p);
        c->cp_field.clear();
        c->cp_state.clear();
        const double need = (double)ncp * (2.0 * (double)c->esz * (double)c->N +
                                           (double)state_total(c));
        size_t free_b = 0, total_b = 0;
        HIPC(c, hipMemGetInfo(&free_b, &total_b));
        if (need > (double)free_b)
            return fail(c, VIRS_ERR_NOMEM, "%d checkpoints need %.1f GB, %.1f GB of device memory are free",
                        ncp, need / 1e9, (double)free_b / 1e9);
        c->cp_field.assign((size_t)ncp * 2, nullptr);
        c->cp_state.assign((size_t)ncp, nullptr);
        for (int k = 0; k < ncp && !rc; ++k) {
            rc = dalloc(c, &c->cp_field[2 * k], c->esz * (size_t)c->N);
            if (!rc) rc = dalloc(c, &c->cp_field[2 * k + 1], c->esz * (size_t)c->N);
            if (!rc) rc = dalloc(c, &c->cp_state[k], state_total(c));
        }
        if (rc) {
            for (auto &p : c->cp_field) dfree(p);
            for (auto &p : c->cp_state) dfree(p);
            c->cp_field.clear();
            c->cp_state.clear();
            return rc;
        }
    }
    c->seg = seg;
    c->ncp = ncp;
    rc = ensure_trace(c, c->seg + 2, ncp == 1 ? 1 : 2);
    if (rc) return rc;
    EarlyStopPause es(c);
    rc = es.pause();
    if (!rc) rc = reset_run(c);
    if (rc) return rc;
    const long long last_base = (long long)(ncp - 1) * c->seg;
    const int last_buf = (ncp - 1) & 1;
    // segments before the last: a checkpoint, then plain steps (graphed like
    // virs_step's), the last two of them writing the filter's d/y history,
    // which the next snapshot keeps for the first steps of its trace; the
    // last segment records its boundary trace step by step
    for (int k = 0; k + 1 < ncp; ++k) {
        rc = save_checkpoint(c, k);
        const long long plain = std::max(0LL, c->seg - 2);
        if (!rc) rc = run_steps(c, plain);
        for (long long s = plain; s < c->seg && !rc; ++s) {
            rc = launch_step(c, nullptr, nullptr, 0, true, nullptr, true);
            if (!rc) ++c->steps;
        }
        if (rc) return rc;
    }
    rc = save_checkpoint(c, ncp - 1);
    if (!rc) rc = seed_trace(c, last_buf, last_base);
    if (rc) return rc;
    for (long long s = last_base; s < nsteps; ++s) {
        rc = launch_step(c, c->trace_d[last_buf], c->trace_y[last_buf], last_base, true);
        if (rc) return rc;
        ++c->steps;
    }
    c->counted = c->steps;
    c->rec_steps = nsteps;
    rc = es.resume();
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return VIRS_OK;
}


// Reverse steps of segment k on the current launch stream.
static int reverse_segment(virs_ctx *c, int k, int &ap, int &aq, int &ar)
{
    const long long T = c->rec_steps;
    const long long s0 = (long long)k * c->seg, s1 = std::min(T, s0 + c->seg);
    const int buf = k & 1;
    for (long long s = s1 - 1; s >= s0; --s) {
        DenseParams p = dense_params(c, c->af[ap], c->af[aq], c->af[ar], c->ctl_adj, ap);
        p.nsoft = 0;
        HIPC(c, launch_dense(c, p, true));
        TailParams t = tail_params(c, c->af[ap], c->af[aq], c->af[ar], c->ctl_adj, ap, ar);
        t.bs = boundary_state(c);
        // the reverse rings: dbar and ybar (FP64)
        t.bs.ws[0] = t.bs.ws[1] = nullptr;
        t.bs.xs[0] = c->ast[0];
        t.bs.xs[1] = c->ast[1];
        t.bs.ys[0] = c->ast[2];
        t.bs.ys[1] = c->ast[3];
        t.bs.trace_d = c->trace_d[buf];
        t.bs.trace_y = c->trace_y[buf];
        t.bs.trace_base = s0;
        t.bs.trace_slots = c->trace_slots;
        t.ir = nullptr;
        t.gsig = c->gsig;
        t.ir_cap = T;
        t.nsoft = 0;
        t.grad_part = c->grad_part;
        HIPC(c, launch_tail_adj(c, t));
        const int tp = ap;
        ap = ar;
        ar = aq;
        aq = tp;
    }
    return VIRS_OK;
}

int virs_adjoint(virs_ctx *c, const float *dl_dir, double *grad)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!grad || (c->nrcv && !dl_dir)) return VIRS_ERR_ARG;
    if (c->rec_steps <= 0 || c->ncp <= 0) return fail(c, VIRS_ERR_STATE, "virs_forward_record first");
    if (c->nmat > 8) return fail(c, VIRS_ERR_UNSUPPORTED, "adjoint supports up to 8 materials");
    const long long T = c->rec_steps;
    // adjoint buffers, second stream and its events
    for (int k = 0; k < 3 && !rc; ++k)
        if (!c->af[k]) rc = dalloc(c, &c->af[k], c->esz * (size_t)c->N);
    for (int k = 0; k < 4 && !rc; ++k)
        if (!c->ast[k]) rc = dalloc(c, &c->ast[k], sizeof(double) * std::max<long long>(1, c->nb));
    if (rc) return rc;
    if (!c->stream2) HIPC(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        if (!c->ev_rec[k]) HIPC(c, hipEventCreateWithFlags(&c->ev_rec[k], hipEventDisableTiming));
        if (!c->ev_rev[k]) HIPC(c, hipEventCreateWithFlags(&c->ev_rev[k], hipEventDisableTiming));
    }
    dfree(c->gsig);
    rc = dalloc(c, &c->gsig, sizeof(double) * std::max(1, c->nrcv) * T);
    if (rc) return rc;
    {
        std::vector<double> g((size_t)std::max(1, c->nrcv) * T, 0.0);
        for (long long i = 0; i < (long long)c->nrcv * T; ++i) g[i] = dl_dir[i];
        HIPC(c, hipMemcpy(c->gsig, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice));
    }
    // one row of partial sums per wave of a list block
    const long long gb = std::max<long long>(1, list_blocks(c->nb, 1, kTailAdjBlock)) * kTailAdjWaves;
    if (c->grad_blocks != gb || !c->grad_part) {
        dfree(c->grad_part);
        rc = dalloc(c, &c->grad_part, sizeof(double) * 5 * 8 * gb);
        if (rc) return rc;
        c->grad_blocks = gb;
    }
    HIPC(c, hipMemsetAsync(c->grad_part, 0, sizeof(double) * 5 * 8 * gb, c->stream));
    rc = zero_fields(c, c->af);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k)
        if (c->nb) HIPC(c, hipMemsetAsync(c->ast[k], 0, sizeof(double) * c->nb, c->stream));
    hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(64), 0, c->stream, c->ctl_adj, 0, T - 1, 1);
    HIPC(c, hipGetLastError());
    EarlyStopPause es(c, c->stream2);
    rc = es.pause();
    if (rc) return rc;
    // the second stream starts after everything queued so far
    HIPC(c, hipStreamSynchronize(c->stream));
    int ap = 0, aq = 1, ar = 2;
    const int last = c->ncp - 1;
    if (last >= 1) {
        OnStream on(c, c->stream2);
        rc = recompute_segment(c, last - 1);
        if (!rc) HIPC(c, hipEventRecord(c->ev_rec[(last - 1) & 1], c->stream));
        if (rc) return rc;
    }
    for (int k = last; k >= 0; --k) {
        if (k != last) HIPC(c, hipStreamWaitEvent(c->stream, c->ev_rec[k & 1], 0));
        rc = reverse_segment(c, k, ap, aq, ar);
        if (rc) return rc;
        HIPC(c, hipEventRecord(c->ev_rev[k & 1], c->stream));
        if (k - 2 >= 0) {
            // segment k-2 reuses the buffer the sweep just enqueued reads
            OnStream on(c, c->stream2);
            HIPC(c, hipStreamWaitEvent(c->stream, c->ev_rev[k & 1], 0));
            rc = recompute_segment(c, k - 2);
            if (!rc) HIPC(c, hipEventRecord(c->ev_rec[(k - 2) & 1], c->stream));
            if (rc) return rc;
        }
    }
    std::vector<double> part((size_t)5 * 8 * gb);
    HIPC(c, hipMemcpyAsync(part.data(), c->grad_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipStreamSynchronize(c->stream2));
    for (int i = 0; i < 5 * c->nmat; ++i) grad[i] = 0.0;
    for (long long b = 0; b < gb; ++b)
        for (int i = 0; i < 5 * c->nmat; ++i) grad[i] += part[(size_t)b * 5 * c->nmat + i];
    // A new virs_forward_record is needed before the next adjoint.  With
    // more than one segment the recomputation left the fields, filter state
    // and step counter at the start of the second segment, so stepping and
    // field reads wait for a reset; with one segment nothing was recomputed
    // and the forward state is the finished run's.
    c->rec_steps = 0;
    c->restart_needed = c->ncp >= 2;
    rc = es.resume();
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return VIRS_OK;
}

} // extern "C"
