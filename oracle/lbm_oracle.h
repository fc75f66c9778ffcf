/*
 * lbm_oracle.h -- CPU FP64 restatement of the reference lattice Boltzmann hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, the smoke check and the cpu_baseline leg of bench.py
 * may load this library to compare the product against it; the product (the HIP library
 * under cylinder-flow_amd/) never links, loads or calls it.
 *
 * Two boundary modes are restated:
 *
 *  "ref"   -- the reference algorithm as written, on the reference's own data layout
 *             (ghost-padded AoS, population index fastest) and in its kernel order:
 *             Initialize, then per step Collision -> boundary fills -> usrbc ->
 *             [momentum exchange] -> stream1 -> stream2 -> Update.
 *               D3Q27: 3D MPI+KOKKOS+MPIIO/lbm.cpp:36-417, lbm.hpp:146-219 (ghost 3)
 *               D2Q9 : MPI Kokkos/lbm.cpp:47-139,313-457, lbm.hpp:77-85     (ghost 2)
 *               D2Q9, single-node variant: Kokkos/lbm.cpp:53-218 (cfg.variant = 1)
 *  "plain" -- the configuration physics (no reference counterpart): half-way bounce-back
 *             on walls and on the solid mask, Zou-He velocity inlet regularised at the
 *             inlet node (Latt & Chopard's projection: plain Zou-He is unstable with BGK
 *             at tau near 1/2), zero-gradient
 *             outlet, optional periodic axes, optional per-node partial bounce-back
 *             (solid fraction s), with the reference's BGK collision and equilibrium.
 *             Written in PUSH form (sources scatter), independent in structure from the
 *             pull-form GPU kernel.
 *
 * Populations are the reference's shifted populations f_i = f_i^std - w_i*rho0, with
 * p = sum_i f_i / 3 and u = sum_i f_i e_i (3D MPI+KOKKOS+MPIIO/lbm.cpp:384-397).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BODY_NONE = 0, ORC_BODY_SPHERE = 1 /* disk in 2D */, ORC_BODY_ZCYL = 2, ORC_BODY_MASK = 3 };
enum { ORC_INLET_UNIFORM = 0, ORC_INLET_PARABOLIC = 1, ORC_INLET_LEGENDRE = 2, ORC_INLET_PROFILE = 3 };

typedef struct orc_cfg {
    int Q;              /* 9 or 27 */
    int nx, ny, nz;     /* physical nodes (nz = 1 for D2Q9) */
    double tau0;        /* reference tau0 = 3*nu; BGK relaxes with 1/(tau0 + 0.5) */
    double u0;          /* inlet velocity scale */
    int body;           /* ORC_BODY_* */
    double cx, cy, cz, r2;  /* body centre and squared radius, in physical node coordinates */
    int inlet;          /* plain mode: ORC_INLET_* */
    int ncoef;
    double coef[8];     /* Legendre coefficients (ORC_INLET_LEGENDRE) */
    uint64_t seed;
    double perturb;     /* plain mode: u_x += perturb * xi, xi in [-1,1) */
    int xper, yper, zper;   /* plain mode: periodic axes (x periodic replaces inlet/outlet) */
    int force_interval; /* record forces every n steps (0: never) */
    int variant;        /* ref mode, D2Q9: 0 the MPI variant (MPI Kokkos/), 1 the single-node
                           variant (Kokkos/lbm.cpp): perturbation 1 + ran*0.001 and bottom/top
                           wall fills over i in [2, lx-2) only */
} orc_cfg;

/* ---- counter-based RNG shared (by specification, not by code) with the GPU init ---- */
uint64_t orc_hash(uint64_t seed, uint64_t index);
double   orc_xi(uint64_t seed, uint64_t index);      /* uniform in [-1, 1) */
int      orc_ran100(uint64_t seed, uint64_t index);  /* integer in [0, 100) */

/* ---- lattice tables in the reference order ---- */
int orc_lattice(int Q, int *e /* Q*3 */, double *w /* Q */, int *bb /* Q */);
/* the reference equilibrium for one node (Q values) */
int orc_feq(int Q, double p, double u, double v, double w, double *out);

/* ---- reference mode ---- */
typedef struct orc_ref orc_ref;
orc_ref *orc_ref_create(const orc_cfg *cfg, const int *ran /* lx*ly*lz padded, or NULL */);
void     orc_ref_destroy(orc_ref *s);
void     orc_ref_extents(const orc_ref *s, int *lx, int *ly, int *lz, int *ghost);
void     orc_ref_step(orc_ref *s, int nsteps);
/* one stage of a step: 0 Collision (step count + 1), 1 fills + usrbc (+ momentum exchange on
 * force steps), 2 stream1 + stream2, 3 Update */
void     orc_ref_stage(orc_ref *s, int stage);
/* fields of the last Update (interior, x fastest) -- w may be NULL in 2D */
void     orc_ref_fields(const orc_ref *s, double *p, double *u, double *v, double *w);
/* populations f at interior nodes, SoA [q][z][y][x] */
void     orc_ref_populations(const orc_ref *s, double *out);
/* whole padded f array in the reference layout (AoS, ii fastest) */
const double *orc_ref_raw(const orc_ref *s);
/* fill padded f with arbitrary values (provenance tracing); ft untouched */
void     orc_ref_set_raw(orc_ref *s, const double *f);
/* the individual reference kernels, for replay tests */
void     orc_ref_collision(orc_ref *s);
void     orc_ref_boundary_fills(orc_ref *s);
void     orc_ref_usrbc(orc_ref *s);
void     orc_ref_mem(orc_ref *s, double F[3]);
void     orc_ref_stream(orc_ref *s);
void     orc_ref_update(orc_ref *s);
int      orc_ref_forces(const orc_ref *s, double *out /* 4 per record */, int max);

/* ---- plain mode ---- */
typedef struct orc_plain orc_plain;
orc_plain *orc_plain_create(const orc_cfg *cfg, const unsigned char *mask /* N or NULL */,
                            const double *sfrac /* N or NULL */, const double *xi /* N or NULL */,
                            const double *uin /* ny*nz or NULL */);
void     orc_plain_destroy(orc_plain *s);
void     orc_plain_step(orc_plain *s, int nsteps);
void     orc_plain_fields(const orc_plain *s, double *p, double *u, double *v, double *w);
void     orc_plain_populations(const orc_plain *s, double *out);   /* SoA [q][z][y][x] */
void     orc_plain_set_populations(orc_plain *s, const double *in);
void     orc_plain_inlet(const orc_plain *s, double *uin);          /* ny*nz */
int      orc_plain_forces(const orc_plain *s, double *out /* 4 per record */, int max);
/* one step's pre-collision populations g (after streaming + boundaries), SoA */
void     orc_plain_last_g(const orc_plain *s, double *out);
int      orc_num_threads(void);
void     orc_set_threads(int n);   /* OpenMP threads of the following calls (n > 0) */

/* ---- adjoint of the plain-mode step (FP64, partial bounce-back body) ---- */
/* J = mean over steps n in [n0, n1] (1-based) of cd_scale * Fx_n, where Fx_n is the
 * momentum-exchange drag of step n.  Returns J; writes dJ/ds (N) and dJ/du_in (ny*nz). */
double   orc_adjoint(const orc_cfg *cfg, const unsigned char *mask, const double *sfrac,
                     const double *xi, const double *uin, int nsteps, int n0, int n1,
                     double cd_scale, int checkpoint_every, double *dJ_ds, double *dJ_duin);
/* forward only: the same objective J (for finite differences) */
double   orc_objective(const orc_cfg *cfg, const unsigned char *mask, const double *sfrac,
                       const double *xi, const double *uin, int nsteps, int n0, int n1,
                       double cd_scale);

#ifdef __cplusplus
}
#endif
#endif
