// Test double of the reference's VectorSpace/Grid (the fields the dispatch
// adapter touches, with the reference's names and types; vectorspace.h:46-105,
// 178-231) and a driver that makes Simulator's calls through the adapter in
// Simulator's order (simulator.cpp:341, 469-474, 569, 571-635, 955-1027).
// Built by the package build into tests/adapter/_build/libdispatch_driver.so;
// the GPU tests call virs_drive_reference_sequence through ctypes.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

struct Grid {
    float *p_curr = nullptr, *p_next = nullptr, *p_prev = nullptr;
    uint8_t *flags = nullptr;
    uint8_t *normals = nullptr;
    size_t num_materials = 0;
    float *biquad_coeffs = nullptr;
    size_t num_filter_sections = 0;
    float *biquad_state_ptr = nullptr;
    bool allocated_filter_memory = false;
    uint32_t *boundary_indices = nullptr;
    size_t boundary_indices_size = 0;
    std::size_t Nx = 0, Ny = 0, Nz = 0, size = 0;
    float *p_audio_output = nullptr;
    size_t p_audio_output_size = 0;
    float percent_to_target_RMS = 0.f;
    inline size_t idx(size_t x, size_t y, size_t z) { return z * Ny * Nx + y * Nx + x; }
};

class VectorSpace {
public:
    VectorSpace(size_t x, size_t y, size_t z, float h) : h(h)
    {
        grid.Nx = x;
        grid.Ny = y;
        grid.Nz = z;
        grid.size = x * y * z;
    }
    Grid &getGrid() { return grid; }
    float h;

private:
    Grid grid;
};

#include "virs_dispatch.hpp"

// the 16 declarations of simulator_dispatch.h:4-25
void fdtd_setup(VectorSpace *space);
bool fdtd_step(VectorSpace *space, unsigned int step);
void fdtd_cleanup(VectorSpace *space);
void fdtd_start_simulation(VectorSpace *space, size_t steps);
void initPressureSphere(VectorSpace *space, size_t xpos, size_t ypos, size_t zpos, size_t radius,
                        float pressure, bool init);
void buildSpongeLayer(VectorSpace *space);
void updateAllGridFromGPU(VectorSpace *space);
void updateCurrentGridFromGPU(VectorSpace *space);
void updateGPUFromGrid(VectorSpace *space);
void uploadNormalsToGPU(VectorSpace *space);
void uploadPZetaToGPU(VectorSpace *space);
void uploadBoundaryIndicesToGPU(VectorSpace *space);
void allocFilterStates(VectorSpace *space);
void allocFilterCoeffs(VectorSpace *space, const size_t num_materials);
void uploadFilterCoeffsToGPU(VectorSpace *space);

VIRS_DEFINE_DISPATCH(VectorSpace)

extern "C" {

/* Simulator's sequence on a voxelized grid:
 *   flags_in   u8[N] voxel flags (0 air, 1 wall, 3 hard source) before the shell
 *   coeffs5    the material row (b0 b1 b2 a1 a2) Simulator writes (simulator.cpp:957-964)
 *   sphere     if radius > 0: initPressureSphere(x, y, z, radius, pressure, true)
 *              after fdtd_start_simulation
 * Outputs: flags_out u8[N] after the shell, ir[steps] (p_audio_output),
 * p_curr/p_prev/p_next [N] from updateAllGridFromGPU, steps_run, the last
 * percent_to_target_RMS and the number of list entries.  Returns 0 or -1. */
int virs_drive_reference_sequence(const uint8_t *flags_in, int64_t nx, int64_t ny, int64_t nz,
                                  float h, const float *coeffs5, int64_t steps,
                                  const int64_t *sphere_xyz_r, float sphere_pressure,
                                  uint8_t *flags_out, float *ir, float *p_curr, float *p_prev,
                                  float *p_next, int64_t *steps_run, float *percent,
                                  int64_t *nboundary)
{
    VectorSpace space((size_t)nx, (size_t)ny, (size_t)nz, h);
    Grid &g = space.getGrid();
    g.num_filter_sections = 3;              // simulator.cpp:339
    g.allocated_filter_memory = false;
    fdtd_setup(&space);
    if (!virs_dispatch::context_of(&space)) return -1;
    std::memcpy(g.flags, flags_in, g.size);
    buildSpongeLayer(&space);
    g.boundary_indices_size = 0;
    updateGPUFromGrid(&space);
    const int64_t dims[3] = {nx, ny, nz};
    virs_compute_normals(g.flags, g.normals, dims, 0);
    uploadNormalsToGPU(&space);
    int64_t nb = 0;
    virs_boundary_list(g.flags, dims, nullptr, &nb);
    std::vector<uint32_t> list((size_t)(nb > 0 ? nb : 1));
    virs_boundary_list(g.flags, dims, list.data(), &nb);
    g.boundary_indices = list.data();
    g.boundary_indices_size = (size_t)nb;
    allocFilterStates(&space);
    uploadBoundaryIndicesToGPU(&space);
    allocFilterCoeffs(&space, 1);
    for (int k = 0; k < 5; ++k) g.biquad_coeffs[k] = coeffs5[k];
    uploadFilterCoeffsToGPU(&space);
    fdtd_start_simulation(&space, (size_t)steps);
    if (sphere_xyz_r && sphere_xyz_r[3] > 0)
        initPressureSphere(&space, (size_t)sphere_xyz_r[0], (size_t)sphere_xyz_r[1],
                           (size_t)sphere_xyz_r[2], (size_t)sphere_xyz_r[3], sphere_pressure, true);
    updateCurrentGridFromGPU(&space);
    int64_t run = 0;
    for (int64_t i = 0; i < steps; ++i) {
        const bool keep = fdtd_step(&space, (unsigned)i);
        ++run;
        if (!keep) break;
    }
    updateAllGridFromGPU(&space);
    std::memcpy(flags_out, g.flags, g.size);
    std::memcpy(ir, g.p_audio_output, sizeof(float) * (size_t)run);
    std::memcpy(p_curr, g.p_curr, sizeof(float) * g.size);
    std::memcpy(p_prev, g.p_prev, sizeof(float) * g.size);
    std::memcpy(p_next, g.p_next, sizeof(float) * g.size);
    *steps_run = run;
    *percent = g.percent_to_target_RMS;
    *nboundary = nb;
    fdtd_cleanup(&space);
    const bool freed = !g.p_curr && !g.p_next && !g.p_prev && !g.flags && !g.normals &&
                       !g.p_audio_output && !g.biquad_coeffs && !g.biquad_state_ptr &&
                       !virs_dispatch::context_of(&space);
    return freed ? 0 : -1;
}

} // extern "C"
