/*
 * lbm_capi.h -- C ABI of the MI355X (gfx950) lattice Boltzmann library liblbm_hip.so.
 *
 * The reference has no FFI layer: its boundary is the C++ class surface that main()
 * drives (System + LBM, 3D MPI+KOKKOS+MPIIO/lbm.hpp:106-233, System.hpp:5-20,
 * main.cpp:16-45).  Each entry point below replaces the reference member function(s)
 * named in its comment; the C++ drop-in classes in cylinder-flow_amd/host/ call them.
 *
 * Conventions
 *   - every function returns LBM_OK (0) or a non-zero LBM_ERR_* code and never throws;
 *     lbm_last_error() describes the last failure of a context (or of lbm_create when
 *     passed NULL);
 *   - the context owns all device memory; caller buffers are host memory borrowed for
 *     the call; work is ordered on the context's own HIP stream and entry points that
 *     return host data synchronise it;
 *   - one context per host thread; contexts are independent;
 *   - fields and populations are exchanged in FP64, x fastest: index x + nx*(y + ny*z),
 *     populations SoA with the direction slowest: q*N + node;
 *   - populations are the reference's shifted populations f_i - w_i*rho0
 *     (3D MPI+KOKKOS+MPIIO/lbm.cpp:263-267,390-397): p = sum f / 3, u = sum f e.
 */
#ifndef LBM_CAPI_H
#define LBM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_INVALID 1   /* bad argument or configuration */
#define LBM_ERR_HIP 2       /* HIP runtime error (message in lbm_last_error) */
#define LBM_ERR_NOMEM 3     /* device or host allocation failed */
#define LBM_ERR_STATE 4     /* call out of order (e.g. step before initialise) */
#define LBM_ERR_UNSTABLE 5  /* a recorded force or a population is NaN/Inf, or a force exceeds the limit */

enum lbm_precision { LBM_FP64 = 0, LBM_FP32 = 1, LBM_FP16S = 2, LBM_BF16S = 3 };
/* LBM_BC_REFERENCE: the reference's own boundary rules and kernel order (ghost layers,
 * link-mirror walls and body, ghost-copy inlet with prescribed-velocity relaxation,
 * lateral outlet extrapolation with p = 0); FP64 only.  D2Q9 follows the MPI 2D variant
 * (MPI Kokkos/lbm.cpp).
 * LBM_BC_REFERENCE_SN: the same for the single-node 2D variant (Kokkos/lbm.cpp:53-218), D2Q9
 * and FP64 only: initial perturbation 1 + ran*0.001 and bottom/top wall fills over the
 * columns i in [2, lx-2) (its two inlet-corner pulls read the previous step's wall fill).
 * LBM_BC_PLAIN: half-way bounce-back walls and body, Zou-He velocity inlet,
 * zero-gradient outlet, optional periodic axes and partial bounce-back. */
enum lbm_bc_mode { LBM_BC_REFERENCE = 0, LBM_BC_PLAIN = 1, LBM_BC_REFERENCE_SN = 2 };
enum lbm_body { LBM_BODY_NONE = 0, LBM_BODY_SPHERE = 1 /* disk in 2D */, LBM_BODY_ZCYL = 2,
                LBM_BODY_MASK = 3 /* lbm_set_mask */ };
enum lbm_inlet { LBM_INLET_UNIFORM = 0, LBM_INLET_PARABOLIC = 1, LBM_INLET_LEGENDRE = 2,
                 LBM_INLET_PROFILE = 3 /* lbm_set_inlet_profile */ };

typedef struct lbm_config {
    int lattice;            /* 9 (D2Q9) or 27 (D3Q27) */
    int nx, ny, nz;         /* physical nodes per axis; nz = 1 for D2Q9 */
    double tau0;            /* reference tau0 = 3*nu; BGK relaxation time is tau0 + 0.5 */
    double rho0;            /* reference density of the shift (1) */
    double u0;              /* inlet velocity scale (lattice units) */
    int precision;          /* enum lbm_precision; compute is FP64 for LBM_FP64, else FP32 */
    int bc_mode;            /* enum lbm_bc_mode */
    int periodic[3];        /* plain mode: periodic x (replaces inlet/outlet), y, z (replace walls) */
    int body;               /* enum lbm_body */
    double cx, cy, cz, r2;  /* body centre and squared radius, physical node coordinates */
    int inlet;              /* plain mode: enum lbm_inlet */
    int ncoef;
    double coef[8];         /* Legendre coefficients (LBM_INLET_LEGENDRE) */
    uint64_t seed;          /* perturbation seed (reference pool seed 5374857 by default) */
    double perturb;         /* plain mode: u_x += perturb * xi, xi uniform in [-1, 1) */
    int force_interval;     /* record drag/lift every n steps (0: never) */
    double force_limit;     /* |F| above this is reported unstable (0: only NaN/Inf) */
    double store_uref;      /* FP16/BF16: populations stored relative to feq(0, (uref,0,0)) */
} lbm_config;

typedef struct lbm_info {
    int lattice, nx, ny, nz, pitch;   /* pitch: padded x extent of a device row */
    size_t nodes;                     /* physical nodes (fluid + solid) */
    size_t state_bytes;               /* bytes of one population buffer on the device */
    double bytes_per_update;          /* algorithmic 2*Q*sizeof(storage) */
    long step;                        /* steps taken since initialise */
    int solid_nodes, boundary_nodes;
    int nodes_per_lane;               /* x nodes per lane of the fused kernel (plain mode) */
} lbm_info;

typedef struct lbm_ctx lbm_ctx;

/* defaults of the reference cases: D3Q27 from 3D MPI+KOKKOS+MPIIO/input.in:1-3 and
 * System.cpp:5-23, D2Q9 from MPI Kokkos/System.cpp:3-18 and main.cpp:8-9 */
int lbm_config_default(lbm_config *cfg, int lattice);
/* the velocity set in the reference order (the "t", "e", "bb" views,
 * 3D MPI+KOKKOS+MPIIO/lbm.cpp:55-218, MPI Kokkos/lbm.cpp:52-90); host-only, no GPU needed.
 * e: 3*Q ints, w: Q doubles, opp: Q ints */
int lbm_lattice_table(int lattice, int *e, double *w, int *opp);

/* LBM::LBM(comm, sx, sy, sz, tau, rho0, u0) -- 3D MPI+KOKKOS+MPIIO/lbm.hpp:146 */
int lbm_create(const lbm_config *cfg, lbm_ctx **out);
/* ~LBM (Kokkos views released) */
void lbm_destroy(lbm_ctx *ctx);
/* no counterpart (the reference aborts on every error) */
const char *lbm_last_error(const lbm_ctx *ctx);

/* the body mask "usr" (3D MPI+KOKKOS+MPIIO/lbm.cpp:238): N bytes, non-zero = solid.
 * Call before lbm_initialize; requires body = LBM_BODY_MASK. */
int lbm_set_mask(lbm_ctx *ctx, const uint8_t *solid);
/* per-node solid fraction s in [0,1] (partial bounce-back, plain mode); N doubles */
int lbm_set_solid_fraction(lbm_ctx *ctx, const double *s);
/* inlet velocity u_x(y, z): ny*nz doubles (LBM_INLET_PROFILE) */
int lbm_set_inlet_profile(lbm_ctx *ctx, const double *uin);
/* inject the initial perturbation instead of the seeded one: plain mode xi (N doubles,
 * in [-1,1)); reference mode the "ran" integers in [0,100) over the ghost-padded box
 * (lx*ly*lz, 3D MPI+KOKKOS+MPIIO/lbm.cpp:224-229) */
int lbm_set_perturbation(lbm_ctx *ctx, const double *xi);
int lbm_set_ran(lbm_ctx *ctx, const int32_t *ran, size_t count);

/* LBM::Initialize -- 3D MPI+KOKKOS+MPIIO/lbm.cpp:36-256 (mask, inlet profile with
 * perturbation, f = feq) */
int lbm_initialize(lbm_ctx *ctx);

/* n time steps: the body of main's loop, Collision + pack/exchange/unpack + Streaming +
 * Update (3D MPI+KOKKOS+MPIIO/main.cpp:29-37; lbm.cpp:257-417), one fused kernel per
 * step; asynchronous on the context stream */
int lbm_step(lbm_ctx *ctx, int n);
/* the same, bracketed by HIP events on the context stream; *ms = elapsed milliseconds */
int lbm_step_timed(lbm_ctx *ctx, int n, double *ms);
/* n steps as lbm_step, timed with HIP events on the context stream: *ms_steps = the n steps,
 * *ms_bulk = the sum over the steps of the bulk kernel alone (k_fast_v; k_dense_v in dense mode;
 * k_fused for FP64 storage), between events recorded just before and after its launch.  Plain mode only,
 * 1 <= n <= 100000.  A measurement entry point with no counterpart in the reference. */
int lbm_step_kernel_timed(lbm_ctx *ctx, int n, double *ms_steps, double *ms_bulk);
/* One time step in the reference's stages, callable one at a time and only in this order
 * (reference boundary modes; LBM_ERR_INVALID in the plain mode, which fuses the stages;
 * LBM_ERR_STATE out of order, and lbm_step / lbm_step_timed inside an unfinished step).
 * Four calls leave the same state as lbm_step(ctx, 1); lbm_get_populations returns the
 * populations as the reference's f stands after the last stage, lbm_get_fields the moments of
 * the last Update.  Asynchronous on the context stream.
 *   lbm_collide   LBM::Collision()  (3D MPI+KOKKOS+MPIIO/lbm.cpp:257-272; MPI Kokkos/lbm.cpp:127-139);
 *                 opens the step: every Collision is applied right after the Update
 *                 before it (the first acts on f = feq and changes nothing), so the
 *                 collided populations are already in place
 *   lbm_boundary  the ghost fills and usrbc of LBM::Streaming() (lbm.cpp:278-356; MPI Kokkos/
 *                 lbm.cpp:315-353), with the momentum exchange on force steps (MPI Kokkos/
 *                 lbm.cpp:356-407)
 *   lbm_stream    stream1 + stream2 of LBM::Streaming() (lbm.cpp:359-370; MPI Kokkos/lbm.cpp:408-416)
 *   lbm_update    LBM::Update() (lbm.cpp:374-417; MPI Kokkos/lbm.cpp:420-457), then the next
 *                 step's Collision in place */
int lbm_collide(lbm_ctx *ctx);
int lbm_boundary(lbm_ctx *ctx);
int lbm_stream(lbm_ctx *ctx);
int lbm_update(lbm_ctx *ctx);
/* drains the context stream, checks the forces recorded since the last check (NaN/Inf or above
 * force_limit) and the state itself (a non-finite rest population at any live node), and reports
 * either as LBM_ERR_UNSTABLE; the reference has no failure detection */
int lbm_sync(lbm_ctx *ctx);

/* macroscopic fields after the last step (the arrays Update writes:
 * p, ua, va, wa -- 3D MPI+KOKKOS+MPIIO/lbm.cpp:384-415); any pointer may be NULL.  Steps
 * write only populations: the first read after a step runs that step again from its input,
 * with the moments on (same populations, bit for bit; the cost of one step), which is how
 * the reference's every-step Update becomes an output-step cost (main.cpp:39-44). */
int lbm_get_fields(lbm_ctx *ctx, double *p, double *u, double *v, double *w);
/* lbm_get_fields without allocating: no counterpart in the reference; for callers that cannot
 * allocate (tight serving or monitoring loops, hosts that forbid allocation after set-up).  The
 * fields are staged through one fixed device array of 32 MiB per GPU (held until the process
 * ends), so at most 4 194 304 nodes (LBM_ERR_INVALID above); calls that overlap on one GPU share
 * that array and race -- serialise them. */
int lbm_get_fields_noalloc(lbm_ctx *ctx, double *p, double *u, double *v, double *w);
/* populations (decoded to FP64), Q*N doubles.  Reference mode: the reference's f after
 * Update (streamed, not yet collided: 3D MPI+KOKKOS+MPIIO/lbm.cpp:359-370).  Plain mode:
 * the stored state (post-collision), which lbm_set_populations replaces (restart). */
int lbm_get_populations(lbm_ctx *ctx, double *f);
int lbm_set_populations(lbm_ctx *ctx, const double *f);
/* the step counter (lbm_info.step), which also fixes the force-record phase (steps with
 * step % force_interval == 0): with lbm_get/set_populations a checkpoint and restart, which the
 * reference does not have; bit-identical continuation in FP64 and FP32 storage */
int lbm_set_step(lbm_ctx *ctx, long step);
/* momentum-exchange force records (step, Fx, Fy, Fz): MPI Kokkos/lbm.cpp:356-407.
 * Copies min(max, available) records, sets *count to the number available. */
int lbm_get_forces(lbm_ctx *ctx, double *out, int max, int *count);
int lbm_info_get(lbm_ctx *ctx, lbm_info *info);

/* LBM::MPIoutput(n) -- 3D MPI+KOKKOS+MPIIO/lbm.cpp:419-739 (binary Tecplot "#!TDV112",
 * x, y, z, u, v, w, p in FORTRAN order; 2D: MPI Kokkos/lbm.cpp:459-737, x, y, u, v, p).
 * The device fields are copied into pinned memory on the context stream and written by
 * a host thread; lbm_snapshot_wait joins it. */
int lbm_snapshot_async(lbm_ctx *ctx, const char *path, int n);
int lbm_snapshot_wait(lbm_ctx *ctx);

/* Discrete adjoint (no reference counterpart).  J = scale * mean over steps n in [n0,n1]
 * (1-based, counted from lbm_initialize) of the x force.  Runs the forward from the
 * initial state with a checkpoint every `checkpoint_every` steps, recomputes each segment
 * and sweeps backwards.  Outputs (any may be NULL): *J, dJ/ds (N doubles), dJ/du_in
 * (ny*nz doubles: inlet velocity per inlet node, including its effect on the initial
 * state).  Plain mode only; leaves the context at step n1. */
int lbm_adjoint(lbm_ctx *ctx, int n0, int n1, double scale, int checkpoint_every,
                double *J, double *dJ_ds, double *dJ_duin, double *ms_forward, double *ms_adjoint);

#ifdef __cplusplus
}
#endif
#endif
