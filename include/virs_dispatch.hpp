/*
 * virs_dispatch.hpp -- the reference's dispatch layer
 * (src/simulator/simulator_dispatch.h:4-25, simulator_dispatch.cpp:1-280)
 * implemented on libvirs_hip.so (virs_fdtd.h) and libvirs_host.so
 * (virs_host.h).
 *
 * Header-only and templated on the caller's VectorSpace type, so it compiles
 * against the reference's own vectorspace.h.  One translation unit on the
 * reference side replaces simulator_dispatch.cpp:
 *
 *     #include "simulator_dispatch.h"
 *     #include "virs_dispatch.hpp"
 *     VIRS_DEFINE_DISPATCH(VectorSpace)
 *
 * which defines the 16 free functions with their reference signatures.
 *
 * Grid contract kept for Simulator (simulator.cpp:262-1082):
 *   fdtd_setup          allocates and zeroes the host mirrors p_curr, p_next,
 *                       p_prev, flags and normals (simulator_gpu.cpp:7-33)
 *   buildSpongeLayer    rewrites grid.flags in place with the boundary shell
 *                       (sponge_kernel.cu:66-99), on the host
 *   fdtd_step           one step; returns the RMS rule's "continue"
 *                       (simulator_gpu.cu:303-330) and keeps
 *                       grid.percent_to_target_RMS
 *   updateCurrentGridFromGPU / updateAllGridFromGPU
 *                       fill the host fields and grid.p_audio_output
 *   fdtd_cleanup        frees what setup, start and the filter allocators
 *                       allocated (simulator_gpu.cpp:237-305)
 * Differences, by design:
 *   - device memory belongs to the library; the Grid's d_* pointers stay null;
 *   - steps are queued without host synchronisation, so p_audio_output is
 *     filled when the caller reads the grid back (Simulator does so after its
 *     loop, simulator.cpp:1027) instead of one 4-byte copy per step;
 *   - fdtd_start_simulation zeroes the device fields and filter state (the
 *     reference leaves whatever the fields held);
 *   - errors print the library's message; fdtd_step then exits, as the
 *     reference's CUDA_CHECK does (simulator_gpu.cu:4-19).
 *
 * Environment overrides, read at fdtd_setup:
 *   VIRS_PRECISION  fp32 (default) | fp16 | bf16 | fp64
 *   VIRS_DEVICE     HIP device ordinal (default 0)
 *   VIRS_RECEIVER   "x,y,z" voxel of the receiver (default 334,225,270,
 *                   the reference's fixed read position, simulator_gpu.cu:295)
 *   VIRS_EARLY_STOP 0 disables the RMS rule (default on, as the reference)
 */
#ifndef VIRS_DISPATCH_HPP
#define VIRS_DISPATCH_HPP

#include "virs_fdtd.h"
#include "virs_host.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace virs_dispatch {

struct Options {
    int32_t precision = VIRS_FP32;
    int32_t device = 0;
    int64_t rcv[3] = {334, 225, 270};
    bool early_stop = true;
    float target_rms = 0.0009f; // TARGET_RMS, simulator_gpu.cu:9
};

struct Session {
    virs_ctx *ctx = nullptr;
    Options opt;
    bool has_receiver = false;
};

inline std::mutex &registry_mutex()
{
    static std::mutex m;
    return m;
}

inline std::unordered_map<const void *, Session> &registry()
{
    static std::unordered_map<const void *, Session> r;
    return r;
}

inline Session *find(const void *space)
{
    std::lock_guard<std::mutex> lk(registry_mutex());
    auto it = registry().find(space);
    return it == registry().end() ? nullptr : &it->second;
}

inline Options options_from_env()
{
    Options o;
    if (const char *p = std::getenv("VIRS_PRECISION")) {
        const std::string s(p);
        if (s == "fp16") o.precision = VIRS_FP16;
        else if (s == "bf16") o.precision = VIRS_BF16;
        else if (s == "fp64") o.precision = VIRS_FP64;
        else o.precision = VIRS_FP32;
    }
    if (const char *p = std::getenv("VIRS_DEVICE")) o.device = std::atoi(p);
    if (const char *p = std::getenv("VIRS_RECEIVER")) {
        long long x, y, z;
        if (std::sscanf(p, "%lld,%lld,%lld", &x, &y, &z) == 3) {
            o.rcv[0] = x;
            o.rcv[1] = y;
            o.rcv[2] = z;
        }
    }
    if (const char *p = std::getenv("VIRS_EARLY_STOP")) o.early_stop = std::atoi(p) != 0;
    return o;
}

inline void report(const char *what, virs_ctx *ctx, int rc)
{
    std::fprintf(stderr, "virs: %s failed (%d): %s\n", what, rc,
                 ctx ? virs_last_error(ctx) : "no context (fdtd_setup not called?)");
}

template <class T>
inline void free_array(T *&p)
{
    delete[] p;
    p = nullptr;
}

// ---------------------------------------------------------------------------

template <class Space>
void fdtd_setup(Space *space)
{
    auto &g = space->getGrid();
    g.p_curr = new float[g.size]();
    g.p_next = new float[g.size]();
    g.p_prev = new float[g.size]();
    g.flags = new uint8_t[g.size]();
    g.normals = new uint8_t[g.size]();
    Session s;
    s.opt = options_from_env();
    virs_desc d;
    virs_desc_default(&d);
    d.nx = (int64_t)g.Nx;
    d.ny = (int64_t)g.Ny;
    d.nz = (int64_t)g.Nz;
    d.h = space->h;
    d.precision = s.opt.precision;
    d.device = s.opt.device;
    const int rc = virs_create(&d, &s.ctx);
    if (rc) {
        report("fdtd_setup", nullptr, rc);
        return;
    }
    std::lock_guard<std::mutex> lk(registry_mutex());
    registry()[space] = s;
}

template <class Space>
bool fdtd_step(Space *space, unsigned int step)
{
    Session *s = find(space);
    if (!s || !s->ctx) {
        report("fdtd_step", nullptr, VIRS_ERR_STATE);
        std::exit(EXIT_FAILURE);
    }
    const int64_t done = virs_steps_done(s->ctx);
    if ((int64_t)step != done) {
        std::fprintf(stderr, "virs: fdtd_step(%u) out of sequence: %lld steps done\n", step,
                     (long long)done);
        std::exit(EXIT_FAILURE);
    }
    int32_t keep = 1;
    const int rc = virs_step(s->ctx, 1, &keep);
    if (rc) {
        report("fdtd_step", s->ctx, rc);
        std::exit(EXIT_FAILURE);
    }
    space->getGrid().percent_to_target_RMS = virs_percent_to_target_rms(s->ctx);
    return keep != 0;
}

template <class Space>
void read_receiver(Space *space, Session *s)
{
    auto &g = space->getGrid();
    if (!s->has_receiver || !g.p_audio_output) return;
    const int64_t done = virs_steps_done(s->ctx);
    if (done <= 0) return;
    std::vector<float> ir((size_t)done);
    const int rc = virs_read_receivers(s->ctx, ir.data());
    if (rc) {
        report("reading the receiver", s->ctx, rc);
        return;
    }
    const size_t n = std::min((size_t)done, (size_t)g.p_audio_output_size);
    std::memcpy(g.p_audio_output, ir.data(), sizeof(float) * n);
}

template <class Space>
void updateCurrentGridFromGPU(Space *space)
{
    Session *s = find(space);
    if (!s || !s->ctx) return report("updateCurrentGridFromGPU", nullptr, VIRS_ERR_STATE);
    auto &g = space->getGrid();
    const int rc = virs_read_field(s->ctx, 0, g.p_curr);
    if (rc) report("updateCurrentGridFromGPU", s->ctx, rc);
    read_receiver(space, s);
}

template <class Space>
void updateAllGridFromGPU(Space *space)
{
    Session *s = find(space);
    if (!s || !s->ctx) return report("updateAllGridFromGPU", nullptr, VIRS_ERR_STATE);
    auto &g = space->getGrid();
    float *dst[3] = {g.p_curr, g.p_prev, g.p_next};
    for (int w = 0; w < 3; ++w) {
        const int rc = virs_read_field(s->ctx, w, dst[w]);
        if (rc) return report("updateAllGridFromGPU", s->ctx, rc);
    }
    read_receiver(space, s);
}

template <class Space>
void updateGPUFromGrid(Space *space)
{
    Session *s = find(space);
    if (!s || !s->ctx) return report("updateGPUFromGrid", nullptr, VIRS_ERR_STATE);
    auto &g = space->getGrid();
    int rc = virs_upload_flags(s->ctx, g.flags);
    if (rc) return report("updateGPUFromGrid (flags)", s->ctx, rc);
    std::vector<double> tmp(g.size);
    const float *src[2] = {g.p_curr, g.p_prev};
    for (int w = 0; w < 2; ++w) {
        for (size_t i = 0; i < g.size; ++i) tmp[i] = (double)src[w][i];
        rc = virs_write_field_f64(s->ctx, w, tmp.data());
        if (rc) return report("updateGPUFromGrid (fields)", s->ctx, rc);
    }
}

template <class Space>
void fdtd_cleanup(Space *space)
{
    auto &g = space->getGrid();
    {
        std::lock_guard<std::mutex> lk(registry_mutex());
        auto it = registry().find(space);
        if (it != registry().end()) {
            if (it->second.ctx) virs_destroy(it->second.ctx);
            registry().erase(it);
        }
    }
    free_array(g.p_curr);
    free_array(g.p_next);
    free_array(g.p_prev);
    free_array(g.flags);
    free_array(g.p_audio_output);
    free_array(g.normals);
    if (g.allocated_filter_memory) {
        free_array(g.biquad_coeffs);
        free_array(g.biquad_state_ptr);
        g.allocated_filter_memory = false;
    }
}

template <class Space>
void fdtd_start_simulation(Space *space, size_t steps)
{
    auto &g = space->getGrid();
    g.p_audio_output_size = steps;
    free_array(g.p_audio_output);
    g.p_audio_output = new float[steps]();
    Session *s = find(space);
    if (!s || !s->ctx) return report("fdtd_start_simulation", nullptr, VIRS_ERR_STATE);
    const int64_t *r = s->opt.rcv;
    s->has_receiver = r[0] >= 0 && r[1] >= 0 && r[2] >= 0 && r[0] < (int64_t)g.Nx &&
                      r[1] < (int64_t)g.Ny && r[2] < (int64_t)g.Nz;
    int rc = VIRS_OK;
    if (s->has_receiver) {
        const int64_t idx = (int64_t)g.idx((size_t)r[0], (size_t)r[1], (size_t)r[2]);
        rc = virs_set_receivers(s->ctx, &idx, 1);
    } else {
        std::fprintf(stderr,
                     "virs: receiver voxel (%lld,%lld,%lld) lies outside the %zux%zux%zu grid; "
                     "no impulse response is recorded (set VIRS_RECEIVER=x,y,z)\n",
                     (long long)r[0], (long long)r[1], (long long)r[2], (size_t)g.Nx, (size_t)g.Ny,
                     (size_t)g.Nz);
        rc = virs_set_receivers(s->ctx, nullptr, 0);
    }
    if (!rc) rc = virs_set_early_stop(s->ctx, s->opt.early_stop && s->has_receiver, s->opt.target_rms);
    if (!rc) rc = virs_start_simulation(s->ctx, (int64_t)steps);
    if (rc) report("fdtd_start_simulation", s->ctx, rc);
}

template <class Space>
void initPressureSphere(Space *space, size_t xpos, size_t ypos, size_t zpos, size_t radius,
                        float pressure, bool init)
{
    // initSphereKernel (init_sphere.cu:4-24) on the current field
    Session *s = find(space);
    if (!s || !s->ctx) return report("initPressureSphere", nullptr, VIRS_ERR_STATE);
    auto &g = space->getGrid();
    std::vector<double> p(g.size);
    int rc = virs_read_field_f64(s->ctx, 0, p.data());
    if (rc) return report("initPressureSphere", s->ctx, rc);
    const size_t r2 = radius * radius;
    for (size_t z = 0; z < g.Nz; ++z)
        for (size_t y = 0; y < g.Ny; ++y)
            for (size_t x = 0; x < g.Nx; ++x) {
                const size_t dx = x > xpos ? x - xpos : xpos - x;
                const size_t dy = y > ypos ? y - ypos : ypos - y;
                const size_t dz = z > zpos ? z - zpos : zpos - z;
                const size_t i = (z * g.Ny + y) * g.Nx + x;
                if (dx * dx + dy * dy + dz * dz <= r2) p[i] = (double)pressure;
                else if (init) p[i] = 0.0;
            }
    rc = virs_write_field_f64(s->ctx, 0, p.data());
    if (rc) report("initPressureSphere", s->ctx, rc);
}

template <class Space>
void buildSpongeLayer(Space *space)
{
    auto &g = space->getGrid();
    const int64_t dims[3] = {(int64_t)g.Nx, (int64_t)g.Ny, (int64_t)g.Nz};
    std::vector<uint8_t> out(g.size);
    const int rc = virs_build_shell(g.flags, out.data(), dims, 0);
    if (rc) {
        std::fprintf(stderr, "virs: buildSpongeLayer failed (%d)\n", rc);
        return;
    }
    std::memcpy(g.flags, out.data(), g.size);
}

template <class Space>
void uploadNormalsToGPU(Space *space)
{
    Session *s = find(space);
    if (!s || !s->ctx) return report("uploadNormalsToGPU", nullptr, VIRS_ERR_STATE);
    const int rc = virs_upload_normals(s->ctx, space->getGrid().normals);
    if (rc) report("uploadNormalsToGPU", s->ctx, rc);
}

template <class Space>
void uploadPZetaToGPU(Space *)
{
    // the reference's body is commented out (simulator_dispatch.cpp:111-118)
}

template <class Space>
void uploadBoundaryIndicesToGPU(Space *space)
{
    Session *s = find(space);
    if (!s || !s->ctx) return report("uploadBoundaryIndicesToGPU", nullptr, VIRS_ERR_STATE);
    auto &g = space->getGrid();
    const int rc = virs_upload_boundary(s->ctx, g.boundary_indices, nullptr, (int64_t)g.boundary_indices_size);
    if (rc) report("uploadBoundaryIndicesToGPU", s->ctx, rc);
}

template <class Space>
void allocFilterStates(Space *space)
{
    // host copy as in simulator_dispatch.cpp:132-160; the device state (FP64)
    // is allocated and zeroed by uploadBoundaryIndicesToGPU
    auto &g = space->getGrid();
    free_array(g.biquad_state_ptr);
    g.biquad_state_ptr = new float[g.num_filter_sections * g.boundary_indices_size * 4]();
    g.allocated_filter_memory = true;
}

template <class Space>
void allocFilterCoeffs(Space *space, const size_t num_materials)
{
    auto &g = space->getGrid();
    g.num_materials = num_materials;
    g.biquad_coeffs = new float[num_materials * 5]();
}

template <class Space>
void uploadFilterCoeffsToGPU(Space *space)
{
    Session *s = find(space);
    if (!s || !s->ctx) return report("uploadFilterCoeffsToGPU", nullptr, VIRS_ERR_STATE);
    auto &g = space->getGrid();
    std::vector<double> c(g.num_materials * 5);
    for (size_t i = 0; i < c.size(); ++i) c[i] = (double)g.biquad_coeffs[i];
    const int rc = virs_set_materials(s->ctx, c.data(), (int32_t)g.num_materials);
    if (rc) report("uploadFilterCoeffsToGPU", s->ctx, rc);
}

/* The context behind a VectorSpace, for callers that want the full C ABI
 * (more receivers, soft sources, reading the IR in FP64). */
inline virs_ctx *context_of(const void *space)
{
    Session *s = find(space);
    return s ? s->ctx : nullptr;
}

} // namespace virs_dispatch

#define VIRS_DEFINE_DISPATCH(SPACE)                                                                \
    void fdtd_setup(SPACE *s) { virs_dispatch::fdtd_setup(s); }                                    \
    bool fdtd_step(SPACE *s, unsigned int step) { return virs_dispatch::fdtd_step(s, step); }      \
    void fdtd_cleanup(SPACE *s) { virs_dispatch::fdtd_cleanup(s); }                                \
    void fdtd_start_simulation(SPACE *s, size_t steps) { virs_dispatch::fdtd_start_simulation(s, steps); } \
    void initPressureSphere(SPACE *s, size_t x, size_t y, size_t z, size_t r, float p, bool init)  \
    {                                                                                              \
        virs_dispatch::initPressureSphere(s, x, y, z, r, p, init);                                 \
    }                                                                                              \
    void buildSpongeLayer(SPACE *s) { virs_dispatch::buildSpongeLayer(s); }                        \
    void updateAllGridFromGPU(SPACE *s) { virs_dispatch::updateAllGridFromGPU(s); }                \
    void updateCurrentGridFromGPU(SPACE *s) { virs_dispatch::updateCurrentGridFromGPU(s); }        \
    void updateGPUFromGrid(SPACE *s) { virs_dispatch::updateGPUFromGrid(s); }                      \
    void uploadNormalsToGPU(SPACE *s) { virs_dispatch::uploadNormalsToGPU(s); }                    \
    void uploadPZetaToGPU(SPACE *s) { virs_dispatch::uploadPZetaToGPU(s); }                        \
    void uploadBoundaryIndicesToGPU(SPACE *s) { virs_dispatch::uploadBoundaryIndicesToGPU(s); }    \
    void allocFilterStates(SPACE *s) { virs_dispatch::allocFilterStates(s); }                      \
    void allocFilterCoeffs(SPACE *s, const size_t m) { virs_dispatch::allocFilterCoeffs(s, m); }   \
    void uploadFilterCoeffsToGPU(SPACE *s) { virs_dispatch::uploadFilterCoeffsToGPU(s); }

#endif
