/*
 * fdtd_oracle.c -- CPU restatement of the VIRS FDTD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Nothing in the product (virs_amd/) links or calls it.
 *
 * What it restates (reference file:line):
 *   oracle_run_f32 / oracle_run_f64
 *     dense update      src/simulator/simulator_gpu.cu:38-103   (fdtd_kernel)
 *     boundary biquad   src/simulator/simulator_gpu.cu:105-150  (fdtd_kernel_boundary_biquad)
 *     hard source       src/simulator/simulator_gpu.cu:59-63,214-218
 *     rotate + receiver src/simulator/simulator_gpu.cu:292-300
 *     constants         src/simulator/simulator_gpu.cu:195-198,147
 *   Extensions of the GPU path (no reference counterpart), each reducing to the
 *   reference when switched off: per-voxel material rows, the filter in
 *   double with two values of state (direct form II, FP64; transposed
 *   direct form II with float-stored states in the 16-bit runner), soft
 *   point sources (added after the boundary correction), several
 *   receivers, graded sponge damping.
 *   oracle_rms_continue   src/simulator/simulator_gpu.cu:303-330
 *   oracle_refcpu_steps   src/simulator/simulator_cpu.cpp:55-81 with
 *                         Grid::laplacian7 src/simulator/vectorspace.h:107-134
 *   oracle_adjoint_f64    reverse-mode transpose of oracle_run_f64 (no reference
 *                         counterpart), written as a literal scatter-form
 *                         transpose of the forward code.
 *   oracle_run_s16        the step with FP16/BF16 field storage and one
 *                         power-of-two scale per field (no reference
 *                         counterpart): the GPU path's 16-bit scheme, which
 *                         restates the dense update above in exact arithmetic
 *                         with walls held at zero.
 *   oracle_round16        its storage rounding (nearest, or stochastic from the
 *                         cell/step hash), exposed for the tests.
 *
 * Pinning (tests/test_oracle.py): rigid-box discrete eigenmodes, discrete
 * energy conservation, source/receiver reciprocity, scipy.signal.lfilter for
 * the filter recursion (float direct form I and double direct form II), a
 * single-cell hand derivation of the boundary correction,
 * central finite differences for the adjoint, brute-force numpy steps on tiny
 * grids for the float32 arithmetic order, and a literal-formula check of the
 * RMS stop on synthetic IRs.  The 16-bit runner: FP16 nearest rounding against
 * numpy's float16 cast, every half pattern round-tripping, BF16 nearest as the
 * nearest value by brute force, stochastic rounding exactly unbiased over every
 * random-bit value, inputs scaled by 2^k giving the same stored bits with the
 * exponents shifted by -k, the exactness in float of the products the fused
 * multiply-adds rely on, and the runner tracking the FP64 scheme.
 *
 * Built with -ffp-contract=off so every multiply and add rounds separately.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t nx, ny, nz;
    /* Reference constants, computed by the caller in float exactly as
     * simulator_gpu.cu:195-198 and :147 do: c2dt2 = c*c*dt*dt,
     * ih2 = 1/(h*h), kconst = (c*dt)/h*0.95f.  The FP64 runner promotes these
     * same float values, so FP32 and FP64 runs share one discrete scheme. */
    float c2dt2, ih2, kconst;
    const uint8_t *flags;           /* [N] 0 air, 1 wall, 2 boundary shell, 3 hard source */
    int32_t sponge_w;               /* width of the graded damping layer, 0 = off */
    const float *spg_a, *spg_b, *spg_g; /* [sponge_w], indexed by distance to the domain edge */
    int64_t nb;                     /* boundary list length */
    const uint32_t *bidx;           /* [nb] ascending linear indices of flag-2 cells */
    const uint8_t *bmat;            /* [nb] material row, NULL = row 0 (the reference) */
    int32_t nmat;
    const double *coeffs;           /* [nmat][5] = b0 b1 b2 a1 a2 */
    int32_t state_f32;              /* 1: reference float state + float coefficients */
    const double *hard_sig;         /* flag-3 value per step, NULL = reference pulse */
    int64_t hard_len;
    int32_t nsoft;
    const int64_t *soft_idx;        /* [nsoft] */
    const double *soft_sig;         /* [nsoft][soft_len] */
    int64_t soft_len;
    int32_t nrcv;
    const int64_t *rcv_idx;         /* [nrcv] */
    const double *init_p, *init_q;  /* optional initial p^0 and p^-1, NULL = zero */
    double *state_io;               /* optional [nb][4] filter state, read at the start of an
                                     * FP32/FP64 run and written at its end, so a run can
                                     * continue in chunks (with init_p/init_q and n0): x1 x2 y1 y2
                                     * of the float direct form I (state_f32), or w1 w2 0 0 of
                                     * the double direct form II */
} oracle_desc;

static int g_threads = 1;

void oracle_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int oracle_get_threads(void) { return g_threads; }

/* Hard-source value at step n: simulator_gpu.cu:214-218 (10 for n<10, else 0). */
static double hard_value(const oracle_desc *d, int64_t n)
{
    if (!d->hard_sig)
        return n < 10 ? 10.0 : 0.0;
    return (n >= 0 && n < d->hard_len) ? d->hard_sig[n] : 0.0;
}

static int sponge_dist(const oracle_desc *d, int x, int y, int z)
{
    int m = x;
    if (d->nx - 1 - x < m) m = d->nx - 1 - x;
    if (y < m) m = y;
    if (d->ny - 1 - y < m) m = d->ny - 1 - y;
    if (z < m) m = z;
    if (d->nz - 1 - z < m) m = d->nz - 1 - z;
    return m;
}

/*
 * One generic runner per precision.  T is the field type; the arithmetic of
 * the dense update is done in T, in the reference's operand order:
 *   lap = (V(x+1)+V(x-1)+V(y+1)+V(y-1)+V(z+1)+V(z-1) - 6*pc) * ih2
 *   R   = a*pc - b*Q + g*(c2dt2*lap)        (a,b,g = 2,1,1 outside the sponge)
 * V(n) = pc if flags[n]==1 else P[n]                     (simulator_gpu.cu:70-81)
 * The boundary correction is evaluated in double as the reference's literal
 * 0.1/0.9 promote it (simulator_gpu.cu:147-148).
 * The boundary biquad (simulator_gpu.cu:130-145) runs either as the
 * reference writes it, direct form I with float state and float
 * coefficients (state_f32), or as the GPU path runs it for FP32/FP64
 * fields, direct form II with double state; both forms are the same filter
 * in exact arithmetic.
 */
#define DEFINE_RUNNER(T, SUFFIX)                                                        \
int oracle_run_##SUFFIX(const oracle_desc *d, int64_t n0, int64_t nsteps, double *ir,      \
                        double *final_p, double *final_q, double *trace_d,                 \
                        double *trace_y)                                                   \
{                                                                                         \
    const int nx = d->nx, ny = d->ny, nz = d->nz;                                         \
    const int64_t sxy = (int64_t)nx * ny;                                                 \
    const int64_t N = sxy * nz;                                                           \
    T *P = (T *)calloc((size_t)N, sizeof(T));                                             \
    T *Q = (T *)calloc((size_t)N, sizeof(T));                                             \
    T *R = (T *)calloc((size_t)N, sizeof(T));                                             \
    double *st = (double *)calloc((size_t)(d->nb > 0 ? d->nb : 1) * 4, sizeof(double));   \
    float *stf = (float *)calloc((size_t)(d->nb > 0 ? d->nb : 1) * 4, sizeof(float));     \
    if (!P || !Q || !R || !st || !stf) {                                                  \
        free(P); free(Q); free(R); free(st); free(stf);                                   \
        return -1;                                                                        \
    }                                                                                     \
    if (d->init_p) for (int64_t i = 0; i < N; ++i) P[i] = (T)d->init_p[i];                \
    if (d->init_q) for (int64_t i = 0; i < N; ++i) Q[i] = (T)d->init_q[i];                \
    if (d->state_io)                                                                      \
        for (int64_t k = 0; k < d->nb * 4; ++k) {                                         \
            st[k] = d->state_io[k];                                                       \
            stf[k] = (float)d->state_io[k];                                               \
        }                                                                                 \
    const T c2dt2 = (T)d->c2dt2, ih2 = (T)d->ih2;                                         \
    for (int64_t s = 0; s < nsteps; ++s) {                                                \
        const int64_t n = n0 + s;                                                         \
        const T src = (T)hard_value(d, n);                                                \
        _Pragma("omp parallel for num_threads(g_threads) schedule(static)")               \
        for (int z = 1; z < nz - 1; ++z) {                                                \
            for (int y = 1; y < ny - 1; ++y) {                                            \
                for (int x = 1; x < nx - 1; ++x) {                                        \
                    const int64_t i = (int64_t)z * sxy + (int64_t)y * nx + x;             \
                    const uint8_t f = d->flags[i];                                        \
                    if (f == 1) { R[i] = (T)0; continue; }                                \
                    if (f == 3) { R[i] = src; continue; }                                 \
                    if (f != 0 && f != 2) continue;                                       \
                    const T pc = P[i];                                                    \
                    const T v0 = d->flags[i + 1] == 1 ? pc : P[i + 1];                    \
                    const T v1 = d->flags[i - 1] == 1 ? pc : P[i - 1];                    \
                    const T v2 = d->flags[i + nx] == 1 ? pc : P[i + nx];                  \
                    const T v3 = d->flags[i - nx] == 1 ? pc : P[i - nx];                  \
                    const T v4 = d->flags[i + sxy] == 1 ? pc : P[i + sxy];                \
                    const T v5 = d->flags[i - sxy] == 1 ? pc : P[i - sxy];                \
                    T sum = v0 + v1;                                                      \
                    sum = sum + v2;                                                       \
                    sum = sum + v3;                                                       \
                    sum = sum + v4;                                                       \
                    sum = sum + v5;                                                       \
                    const T lap = (sum - (T)6 * pc) * ih2;                                \
                    T a = (T)2, b = (T)1, g = (T)1;                                       \
                    if (d->sponge_w > 0) {                                                \
                        const int dd = sponge_dist(d, x, y, z);                           \
                        if (dd < d->sponge_w) {                                           \
                            a = (T)d->spg_a[dd]; b = (T)d->spg_b[dd]; g = (T)d->spg_g[dd]; \
                        }                                                                 \
                    }                                                                     \
                    R[i] = (a * pc - b * Q[i]) + g * (c2dt2 * lap);                       \
                }                                                                         \
            }                                                                             \
        }                                                                                 \
        for (int64_t e = 0; e < d->nb; ++e) {                                             \
            const int64_t bi = d->bidx[e];                                                \
            const int m = d->bmat ? d->bmat[e] : 0;                                       \
            const double *c = d->coeffs + 5 * m;                                          \
            const T dif = P[bi] - Q[bi];                                                  \
            double yv;                                                                    \
            if (d->state_f32) {                                                           \
                float *s4 = stf + 4 * e;                                                  \
                const float in = (float)dif;                                              \
                const float o = (float)c[0] * in + (float)c[1] * s4[0] + (float)c[2] * s4[1] \
                                - (float)c[3] * s4[2] - (float)c[4] * s4[3];              \
                s4[1] = s4[0]; s4[0] = in; s4[3] = s4[2]; s4[2] = o;                      \
                yv = (double)o;                                                           \
            } else {                                                                      \
                /* direct form II in double: w = d - a1 w1 - a2 w2,             */       \
                /* y = b0 w + b1 w1 + b2 w2, state (w1, w2) in s4[0..1]         */       \
                double *s4 = st + 4 * e;                                                  \
                const double in = (double)dif;                                            \
                const double w = in - c[3] * s4[0] - c[4] * s4[1];                        \
                const double o = c[0] * w + c[1] * s4[0] + c[2] * s4[1];                  \
                s4[1] = s4[0]; s4[0] = w;                                                 \
                yv = o;                                                                   \
            }                                                                             \
            if (trace_d) trace_d[s * d->nb + e] = (double)dif;                            \
            if (trace_y) trace_y[s * d->nb + e] = yv;                                     \
            R[bi] = (T)((double)R[bi] -                                                   \
                        (double)d->kconst * (yv * 0.1 + (double)dif * 0.9) * (double)0.4f); \
        }                                                                                 \
        for (int k = 0; k < d->nsoft; ++k) {                                              \
            const int64_t si = d->soft_idx[k];                                            \
            const double v = n < d->soft_len ? d->soft_sig[(int64_t)k * d->soft_len + n] : 0.0; \
            R[si] = R[si] + (T)v;                                                         \
        }                                                                                 \
        T *t = Q; Q = P; P = R; R = t;                                                    \
        for (int r = 0; r < d->nrcv; ++r) ir[(int64_t)r * nsteps + s] = (double)P[d->rcv_idx[r]]; \
    }                                                                                     \
    if (final_p) for (int64_t i = 0; i < N; ++i) final_p[i] = (double)P[i];               \
    if (final_q) for (int64_t i = 0; i < N; ++i) final_q[i] = (double)Q[i];               \
    if (d->state_io)                                                                      \
        for (int64_t k = 0; k < d->nb * 4; ++k)                                           \
            d->state_io[k] = d->state_f32 ? (double)stf[k] : st[k];                       \
    free(P); free(Q); free(R); free(st); free(stf);                                       \
    return 0;                                                                             \
}

DEFINE_RUNNER(float, f32)
DEFINE_RUNNER(double, f64)

/*
 * RMS early stop, simulator_gpu.cu:303-330, with the reference's float
 * arithmetic.  ir[0..step] holds the receiver samples so far; *state is
 * Grid::rms_filter_state and *percent Grid::percent_to_target_RMS.
 * Returns 1 to continue, 0 to stop.
 */
int oracle_rms_continue(const float *ir, uint32_t step, float *state, float *percent,
                        float target_rms)
{
    const uint32_t RMS_SIZE = 100, MIN_STEP = 3000;
    if (step > MIN_STEP) {
        float sum = 0.0f;
        for (uint32_t i = step - RMS_SIZE; i < step; ++i)
            sum += ir[i] * ir[i];
        float rms = sqrtf(sum / (float)RMS_SIZE);
        if (step == MIN_STEP + 1) {
            *state = rms;
        } else {
            const float fg = 0.00008f;
            const float vout = (fg * rms + *state) / (fg + 1.f);
            *state = 2.f * vout - *state;
            rms = vout;
        }
        *percent = target_rms / rms;
        if (rms < target_rms)
            return 0;
    }
    return 1;
}

/*
 * The reference's own CPU step (a different scheme from the GPU one):
 * pressure-release walls via Grid::laplacian7 and uniform damping gamma = 5,
 * simulator_cpu.cpp:55-81, vectorspace.h:107-134.  tbb::parallel_for over z
 * becomes an OpenMP loop over z.  Buffers rotate as in :79-80; on return
 * *P_io holds the newest field.
 */
void oracle_refcpu_steps(int32_t nx, int32_t ny, int32_t nz, const uint8_t *flags, float h,
                         int32_t nsteps, float **P_io, float **Q_io, float **R_io)
{
    const float inv_h2 = 1.f / (h * h);
    const float c = 343.f;
    const float dt = (float)(0.5 * (double)h / (double)(c * sqrtf(3.f)));
    const float c2_dt2 = c * c * dt * dt;
    const float gamma = 5.f;
    const float damp = gamma * dt;
    const int64_t sxy = (int64_t)nx * ny;
    float *P = *P_io, *Q = *Q_io, *R = *R_io;
    for (int s = 0; s < nsteps; ++s) {
        _Pragma("omp parallel for num_threads(g_threads) schedule(static)")
        for (int k = 1; k < nz - 1; ++k) {
            for (int j = 1; j < ny - 1; ++j) {
                for (int i = 1; i < nx - 1; ++i) {
                    const int64_t idx = (int64_t)k * sxy + (int64_t)j * nx + i;
                    float lap = 0.0f;
                    if (flags[idx] != 1) {
                        const float pc = P[idx];
                        lap = (P[idx + 1] + P[idx - 1] + P[idx + nx] + P[idx - nx] +
                               P[idx + sxy] + P[idx - sxy] - 6.0f * pc) * inv_h2;
                    }
                    R[idx] = (2.f - damp) * P[idx] - (1.f - damp) * Q[idx] + c2_dt2 * lap;
                }
            }
        }
        float *t = Q; Q = P; P = R; R = t;
    }
    *P_io = P; *Q_io = Q; *R_io = R;
}

/*
 * Reverse-mode gradient of L(IR) with respect to every material's
 * (b0,b1,b2,a1,a2), for the FP64 scheme of oracle_run_f64, from zero initial
 * fields.  dL_dir is [nrcv][nsteps].  grad is [nmat][5], overwritten.
 * Written as the literal transpose of the forward code, statement by
 * statement in reverse, scattering into the adjoints of the inputs.  Keeps
 * the whole forward filter trajectory in memory: small grids only.
 */
int oracle_adjoint_f64(const oracle_desc *d, int64_t nsteps, const double *dL_dir,
                       double *grad)
{
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t sxy = (int64_t)nx * ny, N = sxy * nz, nb = d->nb;
    const double lam = (double)d->c2dt2 * (double)d->ih2; /* g * c2dt2 * (.. * ih2) */
    const double c2dt2 = d->c2dt2, ih2 = d->ih2;
    double *tr_d = (double *)malloc(sizeof(double) * (size_t)(nsteps * (nb > 0 ? nb : 1)));
    double *tr_y = (double *)malloc(sizeof(double) * (size_t)(nsteps * (nb > 0 ? nb : 1)));
    double *ir = (double *)malloc(sizeof(double) * (size_t)(nsteps * (d->nrcv > 0 ? d->nrcv : 1)));
    double *ub_next = (double *)calloc((size_t)N, sizeof(double)); /* adjoint of u^{n+1} */
    double *ub_cur = (double *)calloc((size_t)N, sizeof(double));  /* adjoint of u^{n}   */
    double *ub_prev = (double *)calloc((size_t)N, sizeof(double)); /* adjoint of u^{n-1} */
    double *sadj = (double *)calloc((size_t)(nb > 0 ? nb : 1) * 4, sizeof(double));
    (void)lam;
    if (!tr_d || !tr_y || !ir || !ub_next || !ub_cur || !ub_prev || !sadj) {
        free(tr_d); free(tr_y); free(ir); free(ub_next); free(ub_cur); free(ub_prev); free(sadj);
        return -1;
    }
    if (d->state_f32 || d->init_p || d->init_q) {
        free(tr_d); free(tr_y); free(ir); free(ub_next); free(ub_cur); free(ub_prev); free(sadj);
        return -2;
    }
    oracle_run_f64(d, 0, nsteps, ir, NULL, NULL, tr_d, tr_y);
    memset(grad, 0, sizeof(double) * 5 * (size_t)d->nmat);
    const double kap = (double)d->kconst * (double)0.4f;
    for (int64_t s = nsteps - 1; s >= 0; --s) {
        /* IR[s] = u^{s+1}[rcv] */
        for (int r = 0; r < d->nrcv; ++r)
            ub_next[d->rcv_idx[r]] += dL_dir[(int64_t)r * nsteps + s];
        /* boundary correction and filter, reversed */
        for (int64_t e = nb - 1; e >= 0; --e) {
            const int64_t bi = d->bidx[e];
            const int m = d->bmat ? d->bmat[e] : 0;
            const double *c = d->coeffs + 5 * m;
            const double rb = ub_next[bi];
            /* R_b = R_dense - K*(0.1*y + 0.9*dif)*0.4 */
            double ybar = -0.1 * kap * rb;
            double dbar = -0.9 * kap * rb;
            /* state after step s: (x1,x2,y1,y2) = (dif, x1_old, y, y1_old) */
            double *sa = sadj + 4 * e;
            dbar += sa[0];
            double x1bar = sa[1];
            ybar += sa[2];
            double y1bar = sa[3];
            double x2bar = 0.0, y2bar = 0.0;
            /* y = b0*dif + b1*x1 + b2*x2 - a1*y1 - a2*y2 */
            const double dv = tr_d[s * nb + e];
            const double x1 = s >= 1 ? tr_d[(s - 1) * nb + e] : 0.0;
            const double x2 = s >= 2 ? tr_d[(s - 2) * nb + e] : 0.0;
            const double y1 = s >= 1 ? tr_y[(s - 1) * nb + e] : 0.0;
            const double y2 = s >= 2 ? tr_y[(s - 2) * nb + e] : 0.0;
            grad[5 * m + 0] += ybar * dv;
            grad[5 * m + 1] += ybar * x1;
            grad[5 * m + 2] += ybar * x2;
            grad[5 * m + 3] += -ybar * y1;
            grad[5 * m + 4] += -ybar * y2;
            dbar += c[0] * ybar;
            x1bar += c[1] * ybar;
            x2bar += c[2] * ybar;
            y1bar += -c[3] * ybar;
            y2bar += -c[4] * ybar;
            sa[0] = x1bar; sa[1] = x2bar; sa[2] = y1bar; sa[3] = y2bar;
            /* dif = P[b] - Q[b] */
            ub_cur[bi] += dbar;
            ub_prev[bi] -= dbar;
        }
        /* dense update, reversed: R_i = a*P_i - b*Q_i + g*(c2dt2*((sum V - 6 P_i)*ih2)) */
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    const int64_t i = (int64_t)z * sxy + (int64_t)y * nx + x;
                    const uint8_t f = d->flags[i];
                    if (f != 0 && f != 2) continue;
                    const double rb = ub_next[i];
                    if (rb == 0.0) continue;
                    double a = 2.0, b = 1.0, g = 1.0;
                    if (d->sponge_w > 0) {
                        const int dd = sponge_dist(d, x, y, z);
                        if (dd < d->sponge_w) { a = d->spg_a[dd]; b = d->spg_b[dd]; g = d->spg_g[dd]; }
                    }
                    const double lbar = g * c2dt2 * ih2 * rb; /* adjoint of (sum V - 6 pc) */
                    ub_cur[i] += a * rb - 6.0 * lbar;
                    ub_prev[i] += -b * rb;
                    const int64_t nbr[6] = {i + 1, i - 1, i + nx, i - nx, i + sxy, i - sxy};
                    for (int k = 0; k < 6; ++k) {
                        if (d->flags[nbr[k]] == 1) ub_cur[i] += lbar;
                        else ub_cur[nbr[k]] += lbar;
                    }
                }
        /* shift: the adjoint of u^{s} becomes the one of the next output */
        double *t = ub_next;
        ub_next = ub_cur;
        ub_cur = ub_prev;
        ub_prev = t;
        memset(ub_prev, 0, sizeof(double) * (size_t)N);
    }
    free(tr_d); free(tr_y); free(ir); free(ub_next); free(ub_cur); free(ub_prev); free(sadj);
    return 0;
}

/* ------------------------------------------------------------------------
 * 16-bit storage runner (FP16 or BF16 fields, FP32 arithmetic, the boundary
 * filter in double as transposed direct form II with its two states stored
 * in float), restating the kernels' scaled-storage rules exactly:
 *   - each stored field has a power-of-two exponent e (stored = value*2^e);
 *   - the exponent of step n's field is 14 - k (clamped to [-60, 60]),
 *     est = m*2^k, where
 *     est = 16*max(M(n-1), M(n-2)) + |hard(n)| + sum_k |soft_k(n)| and M(s)
 *     is the largest |value| the stencil computed for step s before storage
 *     rounding (not the boundary-corrected or source-added values);
 *   - storage rounding is nearest-even or stochastic: 16 bits of a
 *     counter-based hash of (cell, step, salt) are added below the 16-bit
 *     format's last place before truncation (mix32 below);
 *   - walls and border cells hold zero, so the stencil sums the six
 *     neighbours as stored and weights the centre by 6 - (wall neighbours),
 *     which is the reference's substitution V(wall) = pc in exact arithmetic,
 *     and multiplies by the one factor c2dt2*ih2; with stochastic rounding
 *     the last two steps are fused multiply-adds (the FP32 rounding differs
 *     from oracle_run_f32's);
 *   - the boundary list runs after the stencil on the stored field, soft
 *     sources after the boundary correction on its stored result.
 * ------------------------------------------------------------------------ */

/* 16 random bits for cell idx at step n: one 32-bit hash per cell pair,
 * low half for the even cell, high half for the odd one. */
static uint32_t mix32(uint64_t idx, int64_t n, uint32_t salt)
{
    const uint64_t pair = idx & ~(uint64_t)1;
    uint32_t h = (uint32_t)pair * 0x9E3779B1u + ((uint32_t)(pair >> 32) + (uint32_t)n) * 0x85EBCA77u + salt;
    h ^= h >> 16;
    h = (h & 0xffffffu) * 0x1B3C6Du; /* 24-bit multiply */
    h ^= h >> 15;
    return (idx & 1) ? (h >> 16) : (h & 0xffffu);
}

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* IEEE binary16 <-> float, round to nearest even. */
static uint16_t f32_to_f16_rne(float f)
{
    const uint32_t u = f2u(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t mag = u & 0x7fffffffu;
    if (mag >= 0x7f800000u) /* inf / nan */
        return (uint16_t)(sign | 0x7c00u | (mag > 0x7f800000u ? 0x200u | ((mag >> 13) & 0x3ffu) : 0u));
    if (mag >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to inf */
    if (mag >= 0x38800000u) { /* normal half */
        const uint32_t lsb = (mag >> 13) & 1u;
        mag += 0xfffu + lsb;
        return (uint16_t)(sign | (((mag >> 13) - (112u << 10)) & 0x7fffu));
    }
    /* subnormal half (or zero): value = m * 2^-24 */
    if (mag < 0x33000000u) return (uint16_t)sign; /* < 2^-25: rounds to 0 */
    const uint32_t e = mag >> 23;                 /* 102..112 */
    const uint32_t m = (mag & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;              /* bits to drop to reach 2^-24 units */
    uint32_t q = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return (uint16_t)(sign | q);
}

static float f16_to_f32(uint16_t hbits)
{
    const uint32_t sign = (uint32_t)(hbits & 0x8000u) << 16;
    const uint32_t e = (hbits >> 10) & 0x1fu, m = hbits & 0x3ffu;
    if (e == 0) return u2f(sign) + (sign ? -1.f : 1.f) * ldexpf((float)m, -24);
    if (e == 31) return u2f(sign | 0x7f800000u | (m << 13));
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

static uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u = f2u(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static float bf16_to_f32(uint16_t b) { return u2f((uint32_t)b << 16); }

/* float -> binary16 rounding toward zero (finite overflow gives the largest
 * finite half, as the hardware's packed round-toward-zero conversion does). */
static uint16_t f32_to_f16_rtz(float f)
{
    const uint32_t u = f2u(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t mag = u & 0x7fffffffu;
    if (mag > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((mag >> 13) & 0x3ffu));
    if (mag == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
    if (mag >= 0x47800000u) return (uint16_t)(sign | 0x7bffu); /* >= 65536 */
    if (mag >= 0x38800000u) return (uint16_t)(sign | (((mag >> 13) - (112u << 10)) & 0x7fffu));
    if (mag < 0x33800000u) return (uint16_t)sign;              /* < 2^-24 */
    const uint32_t e = mag >> 23;
    const uint32_t m = (mag & 0x7fffffu) | 0x800000u;
    return (uint16_t)(sign | (m >> (126u - e)));
}

/* Stochastic rounding: r (16 random bits, masked to the bits below the
 * format's last place: 13 for FP16, 16 for BF16) is added to the float's bit
 * pattern, the low bits are cleared and the result converted toward zero.
 * In the normal range this is exactly the stochastic rounding of the value
 * (the cleared pattern is representable); in FP16's subnormal range the
 * conversion truncates. */
static uint16_t f32_to_f16_sr(float f, uint32_t r)
{
    return f32_to_f16_rtz(u2f((f2u(f) + (r & 0x1fffu)) & 0xffffe000u));
}

static uint16_t f32_to_bf16_sr(float f, uint32_t r)
{
    return (uint16_t)(((f2u(f) + (r & 0xffffu)) & 0xffff0000u) >> 16);
}

/* salts of the three places a cell may be stored within one step */
#define SALT_DENSE 0u
#define SALT_BOUNDARY 0x68E31DA4u
#define SALT_SOURCE 0xB5297A4Du

typedef struct {
    int bf16;   /* 0 FP16, 1 BF16 */
    int sr;     /* 1 stochastic rounding */
} store16;

static float ld16(const store16 *s, uint16_t v) { return s->bf16 ? bf16_to_f32(v) : f16_to_f32(v); }

static uint16_t st16(const store16 *s, float x, uint64_t idx, int64_t n, uint32_t salt)
{
    if (s->sr) {
        const uint32_t r = mix32(idx, n, salt);
        return s->bf16 ? f32_to_bf16_sr(x, r) : f32_to_f16_sr(x, r);
    }
    return s->bf16 ? f32_to_bf16_rne(x) : f32_to_f16_rne(x);
}

/* The storage rounding of the 16-bit runner on n floats: rbits (16 random
 * bits per value) selects stochastic rounding, NULL round-to-nearest-even.
 * out receives the 16-bit patterns, back (may be NULL) their float values. */
void oracle_round16(const float *in, const uint32_t *rbits, int64_t n, int bf16, uint16_t *out,
                    float *back)
{
    const store16 S = {bf16, rbits != NULL};
    for (int64_t i = 0; i < n; ++i) {
        uint16_t v;
        if (rbits) v = bf16 ? f32_to_bf16_sr(in[i], rbits[i]) : f32_to_f16_sr(in[i], rbits[i]);
        else v = bf16 ? f32_to_bf16_rne(in[i]) : f32_to_f16_rne(in[i]);
        out[i] = v;
        if (back) back[i] = ld16(&S, v);
    }
}

static float exp2i_f(int e) { return u2f((uint32_t)(e + 127) << 23); }

int oracle_run_s16(const oracle_desc *d, int bf16, int sr, int64_t nsteps, double *ir,
                   double *final_p, double *final_q, int32_t *exps)
{
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t sxy = (int64_t)nx * ny, N = sxy * nz;
    const store16 S = {bf16, sr};
    uint16_t *P = (uint16_t *)calloc((size_t)N, 2), *Q = (uint16_t *)calloc((size_t)N, 2),
             *R = (uint16_t *)calloc((size_t)N, 2);
    double *st = (double *)calloc((size_t)(d->nb > 0 ? d->nb : 1) * 4, sizeof(double));
    int64_t *bpos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(d->nsoft > 0 ? d->nsoft : 1));
    if (!P || !Q || !R || !st || !bpos || d->init_p || d->init_q) {
        free(P); free(Q); free(R); free(st); free(bpos);
        return -1;
    }
    for (int k = 0; k < d->nsoft; ++k) {
        bpos[k] = -1;
        for (int64_t e = 0; e < d->nb; ++e)
            if ((int64_t)d->bidx[e] == d->soft_idx[k]) bpos[k] = e;
    }
    int ering[4] = {0, 0, 0, 0};
    float mring[4] = {0.f, 0.f, 0.f, 0.f};
    const float c2dt2 = d->c2dt2, ih2 = d->ih2;
    for (int64_t n = 0; n < nsteps; ++n) {
        const float hard = (float)hard_value(d, n);
        const float sP = exp2i_f(-ering[(n - 1) & 3]), sQ = exp2i_f(-ering[(n - 2) & 3]);
        float est = 16.f * fmaxf(mring[(n - 1) & 3], mring[(n - 2) & 3]) + fabsf(hard);
        for (int k = 0; k < d->nsoft; ++k)
            if (n < d->soft_len) est += fabsf((float)d->soft_sig[(int64_t)k * d->soft_len + n]);
        int eR = 0;
        if (est > 0.f && est < 3.0e38f) {
            int k2;
            frexpf(est, &k2);
            eR = 14 - k2;
            if (eR < -60) eR = -60;
            if (eR > 60) eR = 60;
        }
        const float sR = exp2i_f(eR), sRi = exp2i_f(-eR);
        float mx = 0.f;
        _Pragma("omp parallel for num_threads(g_threads) schedule(static) reduction(max:mx)")
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const int64_t i = (int64_t)z * sxy + (int64_t)y * nx + x;
                    float o = 0.f;
                    const int border = x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1;
                    const uint8_t f = d->flags[i];
                    if (!border && f == 3) {
                        o = hard;
                    } else if (!border && (f == 0 || f == 2)) {
                        /* walls and border cells hold zero in 16-bit storage: the
                         * plain neighbour sum with centre weight kw = 6 - walls
                         * is the reference's rule V(wall) = pc */
                        const float pc = ld16(&S, P[i]) * sP;
                        float sum = ld16(&S, P[i + 1]) * sP + ld16(&S, P[i - 1]) * sP;
                        sum = sum + ld16(&S, P[i + nx]) * sP;
                        sum = sum + ld16(&S, P[i - nx]) * sP;
                        sum = sum + ld16(&S, P[i + sxy]) * sP;
                        sum = sum + ld16(&S, P[i - sxy]) * sP;
                        const int walls = (d->flags[i + 1] == 1) + (d->flags[i - 1] == 1) +
                                          (d->flags[i + nx] == 1) + (d->flags[i - nx] == 1) +
                                          (d->flags[i + sxy] == 1) + (d->flags[i - sxy] == 1);
                        /* t = sum - kw*pc, u = a*pc - b*q, R = t*(g*c2dt2*ih2) + u;
                         * with stochastic rounding u and R are fused
                         * multiply-adds (each rounds once), with nearest
                         * rounding every product rounds apart */
                        const float t = sum - (float)(6 - walls) * pc;
                        float a = 2.f, b = 1.f, g = 1.f;
                        if (d->sponge_w > 0) {
                            const int dd = sponge_dist(d, x, y, z);
                            if (dd < d->sponge_w) { a = d->spg_a[dd]; b = d->spg_b[dd]; g = d->spg_g[dd]; }
                        }
                        const float qq = ld16(&S, Q[i]) * sQ;
                        if (sr)
                            o = fmaf(t, g * (c2dt2 * ih2), fmaf(a, pc, -(b * qq)));
                        else
                            o = (a * pc - b * qq) + g * (t * (c2dt2 * ih2));
                    }
                    if (fabsf(o) > mx) mx = fabsf(o);
                    R[i] = st16(&S, o * sR, (uint64_t)i, n, SALT_DENSE);
                }
        for (int64_t e = 0; e < d->nb; ++e) {
            const int64_t bi = d->bidx[e];
            const int m = d->bmat ? d->bmat[e] : 0;
            const double *c = d->coeffs + 5 * m;
            const float dif = ld16(&S, P[bi]) * sP - ld16(&S, Q[bi]) * sQ;
            /* transposed direct form II in double, states s1, s2 stored as
             * float: y = b0 d + s1, s1' = b1 d - a1 y + s2, s2' = b2 d - a2 y;
             * this step's correction uses the double y */
            double *s4 = st + 4 * e;
            const double in = (double)dif;
            const double yv = c[0] * in + s4[0];
            const double s1 = c[1] * in - c[3] * yv + s4[1];
            const double s2 = c[2] * in - c[4] * yv;
            s4[0] = (double)(float)s1;
            s4[1] = (double)(float)s2;
            float r = ld16(&S, R[bi]) * sRi;
            r = (float)((double)r - (double)d->kconst * (yv * 0.1 + in * 0.9) * (double)0.4f);
            R[bi] = st16(&S, r * sR, (uint64_t)bi, n, SALT_BOUNDARY);
        }
        /* soft sources, on boundary cells too, add to the stored value */
        for (int k = 0; k < d->nsoft; ++k) {
            if (n >= d->soft_len) continue;
            const int64_t si = d->soft_idx[k];
            float r = ld16(&S, R[si]) * sRi;
            r = r + (float)d->soft_sig[(int64_t)k * d->soft_len + n];
            R[si] = st16(&S, r * sR, (uint64_t)si, n, SALT_SOURCE);
        }
        ering[n & 3] = eR;
        mring[n & 3] = mx;
        if (exps) exps[n] = eR;
        uint16_t *t = Q; Q = P; P = R; R = t;
        for (int r = 0; r < d->nrcv; ++r)
            ir[(int64_t)r * nsteps + n] = (double)(ld16(&S, P[d->rcv_idx[r]]) * sRi);
    }
    const float eP = exp2i_f(-ering[(nsteps - 1) & 3]), eQ = exp2i_f(-ering[(nsteps - 2) & 3]);
    if (final_p) for (int64_t i = 0; i < N; ++i) final_p[i] = (double)(ld16(&S, P[i]) * eP);
    if (final_q) for (int64_t i = 0; i < N; ++i) final_q[i] = (double)(ld16(&S, Q[i]) * eQ);
    free(P); free(Q); free(R); free(st); free(bpos);
    return 0;
}
