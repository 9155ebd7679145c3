/*
 * virs_fdtd.h -- C ABI of the MI355X (gfx950) FDTD time-stepping library
 * (libvirs_hip.so).
 *
 * The library replaces the reference's GPU backend behind its dispatch layer:
 *   src/simulator/simulator_dispatch.h:4-25   (16 free functions on VectorSpace*)
 *   src/simulator/simulator_gpu.h:4-10        (fdtd_gpu_setup / fdtd_gpu_step / ...)
 * Each entry point below names the reference function it stands in for.
 * Plain pointers and sizes only; the library owns all device memory, the
 * caller owns every host buffer.  Every call returns VIRS_OK (0) or a negative
 * status; virs_last_error() gives the message.  One context = one HIP stream
 * on one device; contexts share nothing, so several may live in one process
 * (the reference keeps process-global __constant__ state,
 * simulator_gpu.cu:21-36,190-212).
 *
 * Grid layout (Grid::idx, src/simulator/vectorspace.h:102-105):
 *   i = z*Ny*Nx + y*Nx + x, x fastest.  Flags u8 (0 air, 1 wall, 2 boundary
 *   shell, 3 hard source), boundary list u32 ascending.
 */
#ifndef VIRS_FDTD_H
#define VIRS_FDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VIRS_OK 0
#define VIRS_ERR_ARG -1
#define VIRS_ERR_HIP -2
#define VIRS_ERR_STATE -3
#define VIRS_ERR_NOMEM -4
#define VIRS_ERR_UNSUPPORTED -5

/* Storage precision of the pressure fields.  FP16/BF16 store p^n and p^{n-1}
 * with a per-field power-of-two scale and compute in FP32; FP64 is the
 * reference-scheme oracle.  The boundary filter is computed in double in
 * every mode, with two values of state per list entry: kept in FP64 with
 * FP32/FP64 fields, stored in FP32 with 16-bit fields (the reference keeps
 * four values of filter state per entry, in float). */
#define VIRS_FP32 0
#define VIRS_FP16 1
#define VIRS_BF16 2
#define VIRS_FP64 3

/* Storage rounding of FP16/BF16 fields (FP32/FP64 always round to nearest). */
#define VIRS_ROUND_NEAREST 0
#define VIRS_ROUND_STOCHASTIC 1

#define VIRS_MAX_SPONGE 64
#define VIRS_MAX_POINTS 64
/* Cells per grid: boundary-list entries are uint32 cell indices, as in the
 * reference's list (simulator.cpp:571-605, simulator_dispatch.cpp:120-130). */
#define VIRS_MAX_CELLS 4294967296ULL

typedef struct virs_ctx virs_ctx;

typedef struct {
    int64_t nx, ny, nz;     /* grid dimensions (cells) */
    float h;                /* cell size in metres */
    float c;                /* speed of sound, 343 in the reference */
    float dt;               /* time step in seconds; 0 = the reference's
                               0.5f*h/(c*sqrtf(3.f)) (simulator_gpu.cu:196) */
    int32_t precision;      /* VIRS_FP32 / FP16 / BF16 / FP64 */
    int32_t sponge_width;   /* graded damping layer at the domain edge, cells; 0 = none */
    float sponge_sigma;     /* damping per step at the outermost cell */
    int32_t device;         /* HIP device ordinal */
    int32_t use_graph;      /* replay batches of steps as HIP graphs */
    int32_t dense_blocks;   /* stencil grid size; 0 = fill the GPU once */
    int32_t fuse_boundary;  /* 1: run the boundary biquad inside the stencil
                               pass when the list is the grid's interior
                               flag-2 cells; 0: a separate list launch */
    int32_t rounding;       /* VIRS_ROUND_*: storage rounding of 16-bit fields */
    int32_t vector_width;   /* cells per thread along x in the stencil (1, 2, 4, 8
                               up to 16 bytes of storage); 0 = widest dividing nx */
} virs_desc;

/* Fills the reference defaults: c = 343, dt = 0, FP32, no sponge, device 0,
 * boundary list as its own launch, stochastic rounding of 16-bit fields. */
int virs_desc_default(virs_desc *d);

/* fdtd_gpu_setup (simulator_gpu.cpp:7-235): allocate the device grid, zeroed.
 * VIRS_ERR_ARG for a dimension below 3 or a grid above VIRS_MAX_CELLS cells. */
int virs_create(const virs_desc *d, virs_ctx **out);
/* fdtd_gpu_cleanup (simulator_gpu.cpp:237-305). */
int virs_destroy(virs_ctx *ctx);
const char *virs_last_error(const virs_ctx *ctx);

/* Reference constants actually used by the kernels (float, computed as
 * simulator_gpu.cu:195-198,147): dt, c*c*dt*dt, 1/(h*h), (c*dt)/h*0.95f. */
int virs_get_constants(const virs_ctx *ctx, float *dt, float *c2dt2, float *ih2, float *kconst);

/* updateGPUFromGrid, flags part (simulator_dispatch.cpp:70-81).  Values > 3
 * are rejected. */
int virs_upload_flags(virs_ctx *ctx, const uint8_t *flags);
/* uploadNormalsToGPU (simulator_dispatch.cpp:102-109).  Kept on the device
 * as in the reference; no kernel of the step reads it. */
int virs_upload_normals(virs_ctx *ctx, const uint8_t *normals);
/* uploadBoundaryIndicesToGPU + allocFilterStates (simulator_dispatch.cpp:120-160).
 * mat may be NULL (every entry uses material row 0, as the reference does).
 * Filter state is zeroed. */
int virs_upload_boundary(virs_ctx *ctx, const uint32_t *idx, const uint8_t *mat, int64_t nb);
/* allocFilterCoeffs + uploadFilterCoeffsToGPU (simulator_dispatch.cpp:162-280):
 * coeffs[nmat][5] = b0, b1, b2, a1, a2 (a0 = 1).  The filter state is kept:
 * set between steps, the new rows continue from the stored two-value state
 * (the reference sets its rows once, before its time loop). */
int virs_set_materials(virs_ctx *ctx, const double *coeffs, int32_t nmat);

/* Flag-3 (hard) source value per step.  sig == NULL restores the reference
 * pulse: 10 for steps 0-9, then 0 (simulator_gpu.cu:214-218). */
int virs_set_hard_source(virs_ctx *ctx, const double *sig, int64_t len);
/* Soft (additive) point sources, no reference counterpart: after the
 * stencil and the boundary correction, p[idx[k]] += sig[k][n].  Cells must
 * be distinct and must update (air, boundary shell or flag 3): a source on a
 * wall or on the domain border is rejected, here or by a later
 * virs_upload_flags. */
int virs_set_soft_sources(virs_ctx *ctx, const int64_t *idx, const double *sig, int32_t nsrc,
                          int64_t len);
/* Receiver voxels.  The reference has one, fixed at (334,225,270)
 * (simulator_gpu.cu:295). */
int virs_set_receivers(virs_ctx *ctx, const int64_t *idx, int32_t nrcv);
/* RMS early stop of simulator_gpu.cu:303-330 on receiver 0; off by default.
 * The device evaluates the rule at the end of every step and skips every
 * step after the one where it fires, also inside graph replays; virs_step
 * reads the verdict once a step past 3000 has run and then leaves the
 * fields at the step right after the one that fired.  It is off while
 * virs_time_steps, virs_forward_record and virs_adjoint run. */
int virs_set_early_stop(virs_ctx *ctx, int32_t enable, float target_rms);

/* fdtd_start_simulation (simulator_dispatch.cpp:93-100): size the receiver
 * buffer for nsteps, zero fields, filter state and step counter. */
int virs_start_simulation(virs_ctx *ctx, int64_t nsteps);
/* nsteps calls of fdtd_gpu_step (simulator_gpu.cu:184-332), launched without
 * host round trips: with use_graph as replays of a 30-step HIP graph and one
 * graph of the remaining steps, otherwise as single launches.  *keep_going
 * (may be NULL) is 0 once the RMS rule fired; the steps after it are not
 * counted in virs_steps_done.  VIRS_ERR_STATE after a virs_adjoint that
 * recomputed segments, until virs_start_simulation. */
int virs_step(virs_ctx *ctx, int64_t nsteps, int32_t *keep_going);
int64_t virs_steps_done(const virs_ctx *ctx);
/* Waits for the context's stream.  VIRS_ERR_STATE once the run diverged: a
 * receiver sample or a corrected boundary value was not finite (every
 * precision), or the 16-bit scale estimate was (FP16/BF16). */
int virs_synchronize(virs_ctx *ctx);
/* Grid::percent_to_target_RMS of the last step (vectorspace.h:98). */
float virs_percent_to_target_rms(const virs_ctx *ctx);

/* Grid::p_audio_output for every receiver: out[r][k] for k < steps_done. */
int virs_read_receivers(virs_ctx *ctx, float *out);
int virs_read_receivers_f64(virs_ctx *ctx, double *out);
/* updateCurrentGridFromGPU / updateAllGridFromGPU (simulator_dispatch.cpp:31-49):
 * which = 0 current field p^n, 1 previous p^{n-1}, 2 the third buffer (p^{n-2},
 * the reference's p_next after its pointer rotation); values in physical units. */
int virs_read_field(virs_ctx *ctx, int32_t which, float *out);
int virs_read_field_f64(virs_ctx *ctx, int32_t which, double *out);
/* Layer snapshot: the part of updateCurrentGridFromGPU that layerToImage
 * (vectorspace.cpp:77-148, called from simulator.cpp:975-1018) reads, i.e.
 * the XZ plane y = layer of the current field p^n, instead of the whole
 * field.  virs_plane_request queues a gather of the plane and its copy into
 * pinned host memory behind the steps already launched and returns at once
 * with a ticket; virs_plane_fetch waits for that snapshot alone and writes
 * out[z*Nx + x] (Nz*Nx floats, the values virs_read_field gives).  At most 4
 * snapshots are outstanding; a fifth request, or a layer outside [0, Ny),
 * fails with VIRS_ERR_ARG. */
int virs_plane_request(virs_ctx *ctx, int64_t layer, int32_t *ticket);
int virs_plane_fetch(virs_ctx *ctx, int32_t ticket, float *out);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the snapshot of virs_plane_request + virs_plane_fetch, taken
 * synchronously behind the steps already queued, with no allocation at all.
 * The plane is gathered into a fixed device staging array of 2^21 floats
 * (8 MiB) that each GPU the library runs on holds until the process ends,
 * then copied into out[z*Nx + x] (Nz*Nx floats, the caller's buffer).
 * VIRS_ERR_ARG for a layer outside [0, Ny) or a plane above 2^21 cells.
 * Every call on one GPU shares that GPU's staging array: calls that overlap
 * on one GPU (other threads, other contexts on the same device) race. */
int virs_plane_read_noalloc(virs_ctx *ctx, int64_t layer, float *out);
/* updateGPUFromGrid, field part: set p^n (which 0) or p^{n-1} (which 1). */
int virs_write_field_f64(virs_ctx *ctx, int32_t which, const double *in);
/* Power-of-two exponent e of a stored field (which as for virs_read_field;
 * stored = value * 2^e); 0 unless FP16/BF16. */
int virs_field_exponent(virs_ctx *ctx, int32_t which, int32_t *e);

/* Device bytes held by the context. */
int virs_memory_bytes(const virs_ctx *ctx, int64_t *bytes);

/* Timing of a batch of steps (virs_time_steps).  No reference counterpart:
 * the reference's loop is timed by the caller around fdtd_step calls. */
typedef struct {
    double total_ms;          /* HIP events around the timed region on the context's stream */
    double dense_ms;          /* mean stencil launch over the sampling pass */
    double boundary_ms;       /* mean boundary/point launch over the sampling pass */
    int64_t sampled_steps;    /* steps of the sampling pass (before the timed region) */
    int64_t graph_launches;   /* graph replays in the timed region */
    int32_t graph_steps;      /* steps of each full graph (0: none replayed) */
    int32_t tail_graph_steps; /* steps of the one shorter graph at the end (0: none) */
    int64_t single_steps;     /* steps launched one by one in the timed region */
} virs_timing;

/* Runs nsample + nsteps steps.  First a sampling pass of nsample single
 * steps whose two launches each record their own start and stop
 * (hipExtLaunchKernel: the kernel's execution, without the dispatch around
 * it), untimed as a whole; then the timed region: nsteps steps exactly as
 * virs_step runs them (with use_graph: replays of a 30-step graph and one
 * graph of the remaining steps, all built before the region starts;
 * otherwise single launches), between two HIP events and nothing else.
 * The RMS stop is off meanwhile. */
int virs_time_steps(virs_ctx *ctx, int64_t nsteps, int64_t nsample, virs_timing *out);

/* Adjoint gradient of an IR loss with respect to every material's
 * (b0,b1,b2,a1,a2); no reference counterpart.
 * virs_forward_record runs nsteps from zero state keeping ncheckpoints
 * snapshots (p^n, p^{n-1}, filter state); the IR is then readable as usual.
 * With ncheckpoints = 1 the whole boundary trace (2 values per list entry
 * and step) is recorded and nothing is recomputed; ncheckpoints = 0 picks
 * that when the trace takes at most a quarter of the free device memory,
 * else ceil(sqrt(nsteps)) checkpoints.
 * Changing the materials, flags, sources or boundary list afterwards
 * drops the recording (virs_adjoint then fails with VIRS_ERR_STATE).
 * virs_adjoint takes dL/dIR [nrcv][nsteps] (nrcv * nsteps floats, nsteps
 * as recorded) and writes grad[nmat][5]; it
 * recomputes each segment forward from its snapshot and runs the
 * time-reversed stencil with the transposed boundary filters.  The gradient
 * does not depend on the number of checkpoints (bit for bit).  FP32/FP64.
 * With two or more checkpoints the recomputation leaves the fields mid-run:
 * virs_step, the field reads and virs_plane_request then fail with
 * VIRS_ERR_STATE until virs_start_simulation (or virs_forward_record); the
 * recorded IR stays readable. */
int virs_forward_record(virs_ctx *ctx, int64_t nsteps, int32_t ncheckpoints);
int virs_adjoint(virs_ctx *ctx, const float *dl_dir, double *grad);

#ifdef __cplusplus
}
#endif

#endif
