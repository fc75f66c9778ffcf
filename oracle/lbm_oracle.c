/*
 * lbm_oracle.c -- CPU FP64 restatement of the reference LBM hot path (see lbm_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and bench.py's cpu_baseline
 * leg, which compare the product against it.  Never linked into the product library.
 *
 * Built with -ffp-contract=off so every expression rounds exactly as written; the
 * reference's pow(x, 2) is written x*x (both compilers fold pow(x, 2.0) to x*x).
 *
 * Built a second time with -DORC_FLOAT (liboracle_f32.so) for the CPU baseline's FP32
 * legs only: the populations, fields and step arithmetic of both modes are then float (force
 * sums stay double), and the adjoint is left out.  The default build is the FP64 oracle.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* the population, field and step arithmetic type */
#ifdef ORC_FLOAT
typedef float oreal;
#else
typedef double oreal;
#endif

/* ------------------------------------------------------------------------------------ */
/* RNG: SplitMix64 finaliser applied twice (seed mixed with a hashed node index).        */
/* The GPU library implements the same published construction independently.           */
/* ------------------------------------------------------------------------------------ */
static uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
uint64_t orc_hash(uint64_t seed, uint64_t index) { return sm64(seed ^ sm64(index)); }
double orc_xi(uint64_t seed, uint64_t index)
{
    return (double)(orc_hash(seed, index) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}
int orc_ran100(uint64_t seed, uint64_t index) { return (int)(orc_hash(seed, index) % 100ull); }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------ */
/* Lattice tables in the reference order.                                                */
/*   D2Q9 : MPI Kokkos/lbm.cpp:52-90      D3Q27: 3D MPI+KOKKOS+MPIIO/lbm.cpp:55-218       */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int Q, D;
    int e[27][3];
    double t[27];
    int bb[27];
} lattice;

static int make_lattice(int Q, lattice *L)
{
    memset(L, 0, sizeof(*L));
    if (Q == 9) {
        static const int ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
        static const int ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
        static const int bb[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
        L->Q = 9; L->D = 2;
        for (int i = 0; i < 9; i++) {
            L->e[i][0] = ex[i]; L->e[i][1] = ey[i]; L->e[i][2] = 0; L->bb[i] = bb[i];
            L->t[i] = (i == 0) ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
        }
        return 0;
    }
    if (Q == 27) {
        /* rest; 6 axis pairs; 12 edge pairs; 8 corner pairs -- opposite is the next index */
        static const int E[27][3] = {
            {0, 0, 0},
            {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
            {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},
            {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
            {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
            {1, 1, 1}, {-1, -1, -1}, {1, -1, 1}, {-1, 1, -1},
            {1, 1, -1}, {-1, -1, 1}, {1, -1, -1}, {-1, 1, 1}};
        L->Q = 27; L->D = 3;
        for (int i = 0; i < 27; i++) {
            for (int d = 0; d < 3; d++) L->e[i][d] = E[i][d];
            L->bb[i] = (i == 0) ? 0 : ((i % 2) ? i + 1 : i - 1);
            L->t[i] = (i == 0) ? 8.0 / 27.0 : (i < 7 ? 2.0 / 27.0 : (i < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
        }
        return 0;
    }
    return -1;
}

int orc_lattice(int Q, int *e, double *w, int *bb)
{
    lattice L;
    if (make_lattice(Q, &L)) return -1;
    for (int i = 0; i < Q; i++) {
        for (int d = 0; d < 3; d++) e[3 * i + d] = L.e[i][d];
        w[i] = L.t[i];
        bb[i] = L.bb[i];
    }
    return 0;
}

/* Incompressible shifted equilibrium with the reference's expression and rounding order:
 *   feq = t*p*3.0 + t*(3.0*edu + 4.5*eu2 - 1.5*udu)
 * 3D MPI+KOKKOS+MPIIO/lbm.cpp:263-267 ; MPI Kokkos/lbm.cpp:131-134 */
static double feq_ref(const lattice *L, int ii, double p, double u, double v, double w)
{
    double edu = (double)L->e[ii][0] * u + (double)L->e[ii][1] * v;
    double udu = u * u + v * v;
    if (L->D == 3) {
        edu = edu + (double)L->e[ii][2] * w;
        udu = udu + w * w;
    }
    double eu2 = edu * edu;
    return L->t[ii] * p * 3.0 + L->t[ii] * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
}

#ifdef ORC_FLOAT
/* feq_ref's expression in the plain-mode type */
static oreal feq_p(const lattice *L, int ii, oreal p, oreal u, oreal v, oreal w)
{
    oreal edu = (oreal)L->e[ii][0] * u + (oreal)L->e[ii][1] * v;
    oreal udu = u * u + v * v;
    if (L->D == 3) {
        edu = edu + (oreal)L->e[ii][2] * w;
        udu = udu + w * w;
    }
    oreal eu2 = edu * edu;
    return (oreal)L->t[ii] * p * (oreal)3 + (oreal)L->t[ii] * ((oreal)3 * edu + (oreal)4.5 * eu2 - (oreal)1.5 * udu);
}
#else
#define feq_p feq_ref
#endif

int orc_feq(int Q, double p, double u, double v, double w, double *out)
{
    lattice L;
    if (make_lattice(Q, &L)) return -1;
    for (int i = 0; i < Q; i++) out[i] = feq_ref(&L, i, p, u, v, w);
    return 0;
}

static void push_force(double **rec, int *n, int *cap, long step, const double F[3])
{
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 1024;
        *rec = (double *)realloc(*rec, (size_t)*cap * 4 * sizeof(double));
    }
    double *r = *rec + 4 * (size_t)*n;
    r[0] = (double)step; r[1] = F[0]; r[2] = F[1]; r[3] = F[2];
    (*n)++;
}

/* ==================================================================================== */
/* Reference mode                                                                        */
/* ==================================================================================== */
struct orc_ref {
    lattice L;
    int g;                      /* ghost width: 3 (3D) / 2 (2D) */
    int glx, gly, glz, lx, ly, lz;
    double tau0, u0;
    oreal *f, *ft;              /* AoS, ii fastest (Kokkos LayoutLeft of f(ii,i,j,k)) */
    oreal *ua, *va, *wa, *p;
    int *usr, *ran;
    long step;
    int force_interval;
    int variant;                /* 2D: 0 MPI variant, 1 single-node variant (orc_cfg.variant) */
    double *frec; int nrec, caprec;
};

/* flat node index; signed so the 2D stream4 fills that step one column past either
 * edge wrap onto the neighbouring row exactly as LayoutLeft indexing does */
#define RN(s, i, j, k) ((long)(i) + (long)(s)->lx * ((long)(j) + (long)(s)->ly * (long)(k)))
#define RF(s, ii, i, j, k) (s)->f[(long)(ii) + (long)(s)->L.Q * RN(s, i, j, k)]
#define RFT(s, ii, i, j, k) (s)->ft[(long)(ii) + (long)(s)->L.Q * RN(s, i, j, k)]

/* parabolic inlet of the reference: u0*4*c*(n-1-c)/(n-1)^2
 * 3D MPI+KOKKOS+MPIIO/lbm.cpp:239,405 (in z) ; MPI Kokkos/lbm.cpp:109,446 (in y) */
static double ref_profile(double u0, int c, int n)
{
    return u0 * 4.0 * (double)c * (double)(n - 1 - c) / ((double)(n - 1) * (double)(n - 1));
}

static int ref_inside(const orc_cfg *c, int D, int x, int y, int z)
{
    if (c->body == ORC_BODY_NONE) return 0;
    double dx = (double)x - c->cx, dy = (double)y - c->cy;
    double r = dx * dx + dy * dy;
    if (D == 3 && c->body == ORC_BODY_SPHERE) {
        double dz = (double)z - c->cz;
        r = r + dz * dz;
    }
    return r <= c->r2;
}

orc_ref *orc_ref_create(const orc_cfg *c, const int *ran_in)
{
    orc_ref *s = (orc_ref *)calloc(1, sizeof(orc_ref));
    if (make_lattice(c->Q, &s->L)) { free(s); return NULL; }
    const int D = s->L.D, Q = s->L.Q;
    s->g = (D == 3) ? 3 : 2;
    s->glx = c->nx; s->gly = c->ny; s->glz = (D == 3) ? c->nz : 1;
    s->lx = s->glx + 2 * s->g;
    s->ly = s->gly + 2 * s->g;
    s->lz = (D == 3) ? s->glz + 2 * s->g : 1;
    s->tau0 = c->tau0; s->u0 = c->u0;
    s->force_interval = c->force_interval;
    s->variant = D == 2 ? c->variant : 0;
    size_t n = (size_t)s->lx * s->ly * s->lz;
    s->f = (oreal *)calloc(n * Q, sizeof(oreal));
    s->ft = (oreal *)calloc(n * Q, sizeof(oreal));
    s->ua = (oreal *)calloc(n, sizeof(oreal));
    s->va = (oreal *)calloc(n, sizeof(oreal));
    s->wa = (oreal *)calloc(n, sizeof(oreal));
    s->p = (oreal *)calloc(n, sizeof(oreal));
    s->usr = (int *)calloc(n, sizeof(int));
    s->ran = (int *)calloc(n, sizeof(int));
    const int g = s->g;
    /* "random" + "initialize"/"initv" + "initf" over all nodes including ghosts:
     * 3D MPI+KOKKOS+MPIIO/lbm.cpp:221-254 ; MPI Kokkos/lbm.cpp:92-124 */
    for (int k = 0; k < s->lz; k++)
        for (int j = 0; j < s->ly; j++)
            for (int i = 0; i < s->lx; i++) {
                long L = RN(s, i, j, k);
                int r = ran_in ? ran_in[L] : orc_ran100(c->seed, (uint64_t)L);
                s->ran[L] = r;
                int x = i - g, y = j - g, z = (D == 3) ? k - g : 0;
                int usr = ref_inside(c, D, x, y, z) ? 0 : 1;
                s->usr[L] = usr;
                s->p[L] = 0.0;
                s->va[L] = 0.0;
                s->wa[L] = 0.0;
                if (D == 3)
                    s->ua[L] = (oreal)(ref_profile(s->u0, z, s->glz) * (double)usr * (1.0 + (double)r * 0.001));
                else if (s->variant == 1)   /* Kokkos/lbm.cpp:123 */
                    s->ua[L] = (oreal)(ref_profile(s->u0, y, s->gly) * (double)usr * (1.0 + (double)r * 0.001));
                else
                    s->ua[L] = (oreal)(ref_profile(s->u0, y, s->gly) * (double)usr * (0.9 + (double)r * 0.002));
                for (int ii = 0; ii < Q; ii++) {
                    RF(s, ii, i, j, k) = (oreal)feq_ref(&s->L, ii, s->p[L], s->ua[L], s->va[L], s->wa[L]);
                    RFT(s, ii, i, j, k) = 0.0;
                }
            }
    return s;
}

void orc_ref_destroy(orc_ref *s)
{
    if (!s) return;
    free(s->f); free(s->ft); free(s->ua); free(s->va); free(s->wa); free(s->p);
    free(s->usr); free(s->ran); free(s->frec); free(s);
}

void orc_ref_extents(const orc_ref *s, int *lx, int *ly, int *lz, int *ghost)
{
    *lx = s->lx; *ly = s->ly; *lz = s->lz; *ghost = s->g;
}

/* interior box [g, l-g) on every axis (z collapses to {0} in 2D) */
#define KLO(s) ((s)->L.D == 3 ? (s)->g : 0)
#define KHI(s) ((s)->L.D == 3 ? (s)->lz - (s)->g : 1)

/* Collision: 3D MPI+KOKKOS+MPIIO/lbm.cpp:257-272 ; MPI Kokkos/lbm.cpp:127-139 */
void orc_ref_collision(orc_ref *s)
{
    const int Q = s->L.Q, g = s->g;
#pragma omp parallel for collapse(2)
    for (int k = KLO(s); k < KHI(s); k++)
        for (int j = g; j < s->ly - g; j++)
            for (int i = g; i < s->lx - g; i++) {
                long L = RN(s, i, j, k);
                for (int ii = 0; ii < Q; ii++) {
                    oreal feq = feq_p(&s->L, ii, s->p[L], s->ua[L], s->va[L], s->wa[L]);
                    RF(s, ii, i, j, k) = RF(s, ii, i, j, k) - (RF(s, ii, i, j, k) - feq) / (oreal)(s->tau0 + 0.5);
                }
            }
}

/* Domain-boundary fills, in the reference order, in place on f.
 *   3D: front, back, bottom, top, inlet, outlet -- 3D MPI+KOKKOS+MPIIO/lbm.cpp:278-347
 *   2D: inlet, outlet, bottom, top              -- MPI Kokkos/lbm.cpp:315-345 */
void orc_ref_boundary_fills(orc_ref *s)
{
    const lattice *L = &s->L;
    const int Q = L->Q, g = s->g, lx = s->lx, ly = s->ly, lz = s->lz;
    if (L->D == 3) {
        const int ls = g, le0 = g + s->glx, le1 = g + s->gly, le2 = g + s->glz;
        /* bcf (y_lo) */
        for (int ii = 0; ii < Q; ii++)
            if (L->e[ii][1] > 0)
                for (int k = g - 1; k < lz - g + 1; k++)
                    for (int i = g - 1; i < lx - g + 1; i++)
                        RF(s, ii, i, ls - 1, k) = RF(s, L->bb[ii], i + 2 * L->e[ii][0], ls + 1, k + 2 * L->e[ii][2]);
        /* bcb (y_hi) */
        for (int ii = 0; ii < Q; ii++)
            if (L->e[ii][1] < 0)
                for (int k = g - 1; k < lz - g + 1; k++)
                    for (int i = g - 1; i < lx - g + 1; i++)
                        RF(s, ii, i, le1, k) = RF(s, L->bb[ii], i + 2 * L->e[ii][0], le1 - 2, k + 2 * L->e[ii][2]);
        /* bcd (z_lo) */
        for (int ii = 0; ii < Q; ii++)
            if (L->e[ii][2] > 0)
                for (int j = g - 1; j < ly - g + 1; j++)
                    for (int i = g - 1; i < lx - g + 1; i++)
                        RF(s, ii, i, j, ls - 1) = RF(s, L->bb[ii], i + 2 * L->e[ii][0], j + 2 * L->e[ii][1], ls + 1);
        /* bcu (z_hi) */
        for (int ii = 0; ii < Q; ii++)
            if (L->e[ii][2] < 0)
                for (int j = g - 1; j < ly - g + 1; j++)
                    for (int i = g - 1; i < lx - g + 1; i++)
                        RF(s, ii, i, j, le2) = RF(s, L->bb[ii], i + 2 * L->e[ii][0], j + 2 * L->e[ii][1], le2 - 2);
        /* bcl (inlet ghost copy) */
        for (int ii = 0; ii < Q; ii++)
            if (L->e[ii][0] > 0)
                for (int k = g - 1; k < lz - g + 1; k++)
                    for (int j = g - 1; j < ly - g + 1; j++)
                        RF(s, ii, g - 1, j, k) = RF(s, ii, g, j, k);
        /* bcr (outlet lateral extrapolation) */
        for (int ii = 0; ii < Q; ii++)
            if (L->e[ii][0] < 0)
                for (int k = g - 1; k < lz - g + 1; k++)
                    for (int j = g - 1; j < ly - g + 1; j++)
                        RF(s, ii, le0, j, k) = RF(s, ii, le0 + L->e[ii][0], j + L->e[ii][1], k + L->e[ii][2]);
    } else {
        /* stream3, inlet */
        for (int j = 1; j < ly - 1; j++) {
            RF(s, 1, 1, j, 0) = RF(s, 1, 2, j, 0);
            RF(s, 5, 1, j, 0) = RF(s, 5, 2, j, 0);
            RF(s, 8, 1, j, 0) = RF(s, 8, 2, j, 0);
        }
        /* stream3, outlet */
        for (int j = 1; j < ly - 1; j++) {
            RF(s, 3, lx - 2, j, 0) = RF(s, 3, lx - 3, j, 0);
            RF(s, 7, lx - 2, j, 0) = RF(s, 7, lx - 3, j - 1, 0);
            RF(s, 6, lx - 2, j, 0) = RF(s, 6, lx - 3, j + 1, 0);
        }
        /* stream4, bottom (i-2 = -1 and i+2 = lx wrap through the flat index); the single-node
         * variant fills i in [2, lx-2) only (Kokkos/lbm.cpp:158-167), leaving the corner
         * columns to stream3 */
        const int i0 = s->variant == 1 ? 2 : 1, i1 = s->variant == 1 ? lx - 2 : lx - 1;
        for (int i = i0; i < i1; i++) {
            RF(s, 2, i, 1, 0) = RF(s, 4, i, 3, 0);
            RF(s, 5, i, 1, 0) = RF(s, 7, i + 2, 3, 0);
            RF(s, 6, i, 1, 0) = RF(s, 8, i - 2, 3, 0);
        }
        /* stream4, top */
        for (int i = i0; i < i1; i++) {
            RF(s, 4, i, ly - 2, 0) = RF(s, 2, i, ly - 4, 0);
            RF(s, 7, i, ly - 2, 0) = RF(s, 5, i - 2, ly - 4, 0);
            RF(s, 8, i, ly - 2, 0) = RF(s, 6, i + 2, ly - 4, 0);
        }
    }
}

/* usrbc: solid X with fluid X+e_i gets f_i(X) <- f_bb(i)(X + 2 e_i), in place.
 * 3D MPI+KOKKOS+MPIIO/lbm.cpp:350-356 ; MPI Kokkos/lbm.cpp:347-353 */
void orc_ref_usrbc(orc_ref *s)
{
    const lattice *L = &s->L;
    const int Q = L->Q, g = s->g;
    for (int k = KLO(s); k < KHI(s); k++)
        for (int j = g; j < s->ly - g; j++)
            for (int i = g; i < s->lx - g; i++) {
                if (s->usr[RN(s, i, j, k)] != 0) continue;
                for (int ii = 0; ii < Q; ii++) {
                    const int *e = L->e[ii];
                    if (s->usr[RN(s, i + e[0], j + e[1], k + e[2])] == 1)
                        RF(s, ii, i, j, k) = RF(s, L->bb[ii], i + 2 * e[0], j + 2 * e[1], k + 2 * e[2]);
                }
            }
}

/* Momentum exchange after usrbc: F = sum over fluid Y, solid Y+e_i of
 * (f_i(Y) + f_bb(i)(Y + e_i)) e_i.  MPI Kokkos/lbm.cpp:364-392 (2D); the same formula is
 * applied in 3D, where the reference has no force.  Summed per node, then nodes x-fastest. */
void orc_ref_mem(orc_ref *s, double F[3])
{
    const lattice *L = &s->L;
    const int Q = L->Q, g = s->g;
    F[0] = F[1] = F[2] = 0.0;
    for (int k = KLO(s); k < KHI(s); k++)
        for (int j = g; j < s->ly - g; j++)
            for (int i = g; i < s->lx - g; i++) {
                if (s->usr[RN(s, i, j, k)] != 1) continue;
                double fx = 0.0, fy = 0.0, fz = 0.0;
                for (int ii = 0; ii < Q; ii++) {
                    const int *e = L->e[ii];
                    if (s->usr[RN(s, i + e[0], j + e[1], k + e[2])] == 0) {
                        double m = RF(s, ii, i, j, k) + RF(s, L->bb[ii], i + e[0], j + e[1], k + e[2]);
                        fx += m * (double)e[0];
                        fy += m * (double)e[1];
                        fz += m * (double)e[2];
                    }
                }
                F[0] += fx; F[1] += fy; F[2] += fz;
            }
}

/* stream1 (pull into ft) + stream2 (copy back), interior only:
 * 3D MPI+KOKKOS+MPIIO/lbm.cpp:359-370 ; MPI Kokkos/lbm.cpp:408-416 */
void orc_ref_stream(orc_ref *s)
{
    const lattice *L = &s->L;
    const int Q = L->Q, g = s->g;
#pragma omp parallel for collapse(2)
    for (int k = KLO(s); k < KHI(s); k++)
        for (int j = g; j < s->ly - g; j++)
            for (int i = g; i < s->lx - g; i++)
                for (int ii = 0; ii < Q; ii++) {
                    const int *e = L->e[ii];
                    RFT(s, ii, i, j, k) = RF(s, ii, i - e[0], j - e[1], k - e[2]);
                }
#pragma omp parallel for collapse(2)
    for (int k = KLO(s); k < KHI(s); k++)
        for (int j = g; j < s->ly - g; j++)
            for (int i = g; i < s->lx - g; i++)
                for (int ii = 0; ii < Q; ii++) RF(s, ii, i, j, k) = RFT(s, ii, i, j, k);
}

/* Update: zero all macros, then the serial moment loop with the inlet velocity and the
 * outlet pressure overrides.  3D MPI+KOKKOS+MPIIO/lbm.cpp:374-417 ; MPI Kokkos/lbm.cpp:420-457 */
void orc_ref_update(orc_ref *s)
{
    const lattice *L = &s->L;
    const int Q = L->Q, g = s->g;
    size_t n = (size_t)s->lx * s->ly * s->lz;
    memset(s->ua, 0, n * sizeof(oreal));
    memset(s->va, 0, n * sizeof(oreal));
    memset(s->wa, 0, n * sizeof(oreal));
    memset(s->p, 0, n * sizeof(oreal));
    for (int k = KLO(s); k < KHI(s); k++)
        for (int j = g; j < s->ly - g; j++)
            for (int i = g; i < s->lx - g; i++) {
                long N = RN(s, i, j, k);
                for (int ii = 0; ii < Q; ii++) {
                    oreal fv = RF(s, ii, i, j, k);
                    s->p[N] = s->p[N] + fv / (oreal)3;
                    s->ua[N] = s->ua[N] + fv * (oreal)L->e[ii][0];
                    s->va[N] = s->va[N] + fv * (oreal)L->e[ii][1];
                    if (L->D == 3) s->wa[N] = s->wa[N] + fv * (oreal)L->e[ii][2];
                }
                oreal us = (oreal)s->usr[N];
                s->ua[N] = s->ua[N] * us;
                s->va[N] = s->va[N] * us;
                s->wa[N] = s->wa[N] * us;
                s->p[N] = s->p[N] * us;
                long Nin = RN(s, g, j, k);
                if (L->D == 3)
                    s->ua[Nin] = (oreal)ref_profile(s->u0, k - g, s->glz);
                else
                    s->ua[Nin] = (oreal)ref_profile(s->u0, j - g, s->gly);
                s->va[Nin] = 0.0;
                s->wa[Nin] = 0.0;
                s->p[RN(s, s->lx - g - 1, j, k)] = 0.0;
            }
}

void orc_ref_step(orc_ref *s, int nsteps)
{
    for (int it = 0; it < nsteps; it++) {
        s->step++;
        orc_ref_collision(s);
        orc_ref_boundary_fills(s);
        orc_ref_usrbc(s);
        if (s->force_interval > 0 && s->step % s->force_interval == 0) {
            double F[3];
            orc_ref_mem(s, F);
            push_force(&s->frec, &s->nrec, &s->caprec, s->step, F);
        }
        orc_ref_stream(s);
        orc_ref_update(s);
    }
}

/* One stage of a step, for the staged C ABI: 0 Collision (opens the step: step count + 1),
 * 1 the fills + usrbc part of Streaming with the momentum exchange on force steps, 2 stream1 +
 * stream2, 3 Update.  The reference's method split: 3D MPI+KOKKOS+MPIIO/lbm.cpp:257-272
 * (Collision), 278-370 (Streaming), 374-417 (Update); main.cpp:29-37 calls them in this order. */
void orc_ref_stage(orc_ref *s, int stage)
{
    switch (stage) {
    case 0:
        s->step++;
        orc_ref_collision(s);
        break;
    case 1:
        orc_ref_boundary_fills(s);
        orc_ref_usrbc(s);
        if (s->force_interval > 0 && s->step % s->force_interval == 0) {
            double F[3];
            orc_ref_mem(s, F);
            push_force(&s->frec, &s->nrec, &s->caprec, s->step, F);
        }
        break;
    case 2:
        orc_ref_stream(s);
        break;
    default:
        orc_ref_update(s);
        break;
    }
}

void orc_ref_fields(const orc_ref *s, double *p, double *u, double *v, double *w)
{
    const int g = s->g, gz = (s->L.D == 3) ? g : 0;
    size_t o = 0;
    for (int k = 0; k < s->glz; k++)
        for (int j = 0; j < s->gly; j++)
            for (int i = 0; i < s->glx; i++, o++) {
                long N = RN(s, i + g, j + g, k + gz);
                if (p) p[o] = s->p[N];
                if (u) u[o] = s->ua[N];
                if (v) v[o] = s->va[N];
                if (w) w[o] = s->wa[N];
            }
}

void orc_ref_populations(const orc_ref *s, double *out)
{
    const int g = s->g, gz = (s->L.D == 3) ? g : 0;
    const size_t N = (size_t)s->glx * s->gly * s->glz;
    for (int ii = 0; ii < s->L.Q; ii++) {
        size_t o = 0;
        for (int k = 0; k < s->glz; k++)
            for (int j = 0; j < s->gly; j++)
                for (int i = 0; i < s->glx; i++, o++) out[ii * N + o] = RF(s, ii, i + g, j + g, k + gz);
    }
}

#ifndef ORC_FLOAT
const double *orc_ref_raw(const orc_ref *s) { return s->f; }
#endif
void orc_ref_set_raw(orc_ref *s, const double *f)
{
    for (size_t k = 0; k < (size_t)s->lx * s->ly * s->lz * s->L.Q; k++) s->f[k] = (oreal)f[k];
}

int orc_ref_forces(const orc_ref *s, double *out, int max)
{
    int n = s->nrec < max ? s->nrec : max;
    if (out) memcpy(out, s->frec, (size_t)n * 4 * sizeof(double));
    return s->nrec;
}

/* ==================================================================================== */
/* Plain mode (push form)                                                                */
/* ==================================================================================== */
enum { C_NORMAL = 0, C_ZH = 1, C_WALL = 2, C_OUTLET = 3, C_MASK = 4 };

struct orc_plain {
    lattice L;
    int nx, ny, nz;
    size_t N;
    double tau0, u0;
    int xper, yper, zper;
    unsigned char *solid;
    double *s;                  /* solid fraction or NULL */
    double *uin;                /* ny*nz */
    oreal *h, *g;               /* SoA [q][N] */
    double *mp, *mu, *mv, *mw;
    double *rowF;               /* 3 per (y,z) row */
    long step;
    int force_interval;
    double *frec; int nrec, caprec;
};

static inline int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

/* source classification of the pulled entry (node (x,y,z), direction i) */
static int classify(const orc_plain *s, int x, int y, int z, int i, size_t *src)
{
    const int *e = s->L.e[i];
    if (!s->xper && x == 0 && e[0] > 0) return C_ZH;
    int sx = x - e[0], sy = y - e[1], sz = z - e[2];
    if (s->yper) sy = wrapi(sy, s->ny);
    else if (sy < 0 || sy >= s->ny) return C_WALL;
    if (s->L.D == 3) {
        if (s->zper) sz = wrapi(sz, s->nz);
        else if (sz < 0 || sz >= s->nz) return C_WALL;
    }
    if (s->xper) sx = wrapi(sx, s->nx);
    else if (sx > s->nx - 1) {
        *src = (size_t)(s->nx - 1) + (size_t)s->nx * ((size_t)sy + (size_t)s->ny * sz);
        return C_OUTLET;
    }
    size_t sn = (size_t)sx + (size_t)s->nx * ((size_t)sy + (size_t)s->ny * sz);
    if (s->solid[sn]) return C_MASK;
    *src = sn;
    return C_NORMAL;
}

static double legendre(int k, double x)
{
    double p0 = 1.0, p1 = x;
    if (k == 0) return p0;
    for (int n = 1; n < k; n++) {
        double p2 = ((2.0 * n + 1.0) * x * p1 - (double)n * p0) / (double)(n + 1);
        p0 = p1; p1 = p2;
    }
    return p1;
}

/* inlet profiles (no reference counterpart: the configurations' Zou-He inlet).  Walls
 * sit half a spacing outside the first/last node, so eta = (y + 0.5)/ny. */
static void plain_inlet(const orc_cfg *c, int D, double *uin)
{
    static const int LY[8] = {0, 1, 0, 2, 1, 0, 3, 0}, LZ[8] = {0, 0, 1, 0, 1, 2, 0, 3};
    const int ny = c->ny, nz = (D == 3) ? c->nz : 1;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++) {
            double ey = ((double)y + 0.5) / (double)ny, ez = ((double)z + 0.5) / (double)nz;
            double by = 4.0 * ey * (1.0 - ey), bz = (D == 3) ? 4.0 * ez * (1.0 - ez) : 1.0;
            double u;
            if (c->inlet == ORC_INLET_PARABOLIC) {
                u = c->u0 * by * bz;
            } else if (c->inlet == ORC_INLET_LEGENDRE) {
                double sum = 0.0;
                for (int k = 0; k < c->ncoef && k < 8; k++) {
                    double phi = (D == 3) ? legendre(LY[k], 2.0 * ey - 1.0) * legendre(LZ[k], 2.0 * ez - 1.0)
                                          : legendre(k, 2.0 * ey - 1.0);
                    sum += c->coef[k] * phi;
                }
                u = by * bz * sum;
            } else {
                u = c->u0;
            }
            uin[y + (size_t)ny * z] = u;
        }
}

orc_plain *orc_plain_create(const orc_cfg *c, const unsigned char *mask, const double *sfrac,
                            const double *xi, const double *uin)
{
    orc_plain *s = (orc_plain *)calloc(1, sizeof(orc_plain));
    if (make_lattice(c->Q, &s->L)) { free(s); return NULL; }
    const int D = s->L.D, Q = s->L.Q;
    s->nx = c->nx; s->ny = c->ny; s->nz = (D == 3) ? c->nz : 1;
    s->N = (size_t)s->nx * s->ny * s->nz;
    s->tau0 = c->tau0; s->u0 = c->u0;
    s->xper = c->xper; s->yper = c->yper; s->zper = (D == 3) ? c->zper : 0;
    s->force_interval = c->force_interval;
    const size_t N = s->N;
    s->solid = (unsigned char *)calloc(N, 1);
    s->uin = (double *)calloc((size_t)s->ny * s->nz, sizeof(double));
    s->h = (oreal *)malloc(N * Q * sizeof(oreal));
    s->g = (oreal *)malloc(N * Q * sizeof(oreal));
    s->mp = (double *)calloc(N, sizeof(double));
    s->mu = (double *)calloc(N, sizeof(double));
    s->mv = (double *)calloc(N, sizeof(double));
    s->mw = (double *)calloc(N, sizeof(double));
    s->rowF = (double *)calloc((size_t)s->ny * s->nz * 3, sizeof(double));
    if (sfrac) {
        s->s = (double *)malloc(N * sizeof(double));
        memcpy(s->s, sfrac, N * sizeof(double));
    }
    if (uin) memcpy(s->uin, uin, (size_t)s->ny * s->nz * sizeof(double));
    else plain_inlet(c, D, s->uin);
#pragma omp parallel for
    for (long n = 0; n < (long)N; n++) {
        int x = (int)(n % s->nx), y = (int)((n / s->nx) % s->ny), z = (int)(n / ((size_t)s->nx * s->ny));
        int sol;
        if (mask) sol = mask[n] != 0;
        else if (c->body == ORC_BODY_MASK || c->body == ORC_BODY_NONE) sol = 0;
        else sol = ref_inside(c, D, x, y, z);
        s->solid[n] = (unsigned char)sol;
        double u = 0.0;
        if (!sol) u = s->uin[y + (size_t)s->ny * z] + c->perturb * (xi ? xi[n] : orc_xi(c->seed, (uint64_t)n));
        for (int ii = 0; ii < Q; ii++) s->h[ii * N + n] = (oreal)(sol ? 0.0 : feq_ref(&s->L, ii, 0.0, u, 0.0, 0.0));
        s->mu[n] = u;
    }
    return s;
}

void orc_plain_destroy(orc_plain *s)
{
    if (!s) return;
    free(s->solid); free(s->uin); free(s->h); free(s->g); free(s->s);
    free(s->mp); free(s->mu); free(s->mv); free(s->mw); free(s->rowF); free(s->frec); free(s);
}

/* Zou-He velocity inlet for the shifted incompressible model: for each unknown i (e_ix>0)
 *   g_i = g_bb(i) + 6 w_i U - 18 w_i (e_iy N_y + e_iz N_z),  N_t = sum_{e_kx=0} g_k e_kt,
 * which reproduces sum_i g_i e_i = (U, 0, 0) exactly (no reference counterpart). */
static void zou_he(const lattice *L, oreal *gv, oreal U)
{
    oreal Ny = 0.0, Nz = 0.0;
    for (int k = 0; k < L->Q; k++) {
        if (L->e[k][0] != 0) continue;
        if (L->e[k][1] > 0) Ny += gv[k]; else if (L->e[k][1] < 0) Ny -= gv[k];
        if (L->e[k][2] > 0) Nz += gv[k]; else if (L->e[k][2] < 0) Nz -= gv[k];
    }
    for (int i = 0; i < L->Q; i++) {
        if (L->e[i][0] <= 0) continue;
        oreal tr = 0.0;
        if (L->e[i][1] > 0) tr += Ny; else if (L->e[i][1] < 0) tr -= Ny;
        if (L->e[i][2] > 0) tr += Nz; else if (L->e[i][2] < 0) tr -= Nz;
        gv[i] = (gv[L->bb[i]] + (oreal)(6.0 * L->t[i]) * U) - (oreal)(18.0 * L->t[i]) * tr;
    }
}

/* Regularised inlet (Latt & Chopard's regularisation of the Zou-He node; no reference
 * counterpart): with S = sum g and P_ab = sum g e_a e_b - S/3 delta_ab of the Zou-He-completed
 * populations,  g_i = w_i (S + 3 e_ix U + 4.5 Q_i:P),  Q_i = e_i e_i - delta/3.  The sums run
 * in direction order and every product is written out, so the FP64 GPU path (built, like
 * this file, without FMA contraction) reproduces it bit for bit. */
static void regularize_inlet(const lattice *L, oreal *g, oreal U)
{
    oreal S = 0.0, Mxx = 0.0, Myy = 0.0, Mzz = 0.0, Mxy = 0.0, Mxz = 0.0, Myz = 0.0;
    for (int k = 0; k < L->Q; k++) {
        const int ex = L->e[k][0], ey = L->e[k][1], ez = L->e[k][2];
        S = S + g[k];
        if (ex != 0) Mxx = Mxx + g[k];
        if (ey != 0) Myy = Myy + g[k];
        if (ez != 0) Mzz = Mzz + g[k];
        if (ex * ey > 0) Mxy = Mxy + g[k]; else if (ex * ey < 0) Mxy = Mxy - g[k];
        if (ex * ez > 0) Mxz = Mxz + g[k]; else if (ex * ez < 0) Mxz = Mxz - g[k];
        if (ey * ez > 0) Myz = Myz + g[k]; else if (ey * ez < 0) Myz = Myz - g[k];
    }
    const oreal S3 = S * (oreal)(1.0 / 3.0);
    const oreal Pxx = Mxx - S3, Pyy = Myy - S3, Pzz = Mzz - S3;
    oreal tr = Pxx + Pyy;
    if (L->D == 3) tr = tr + Pzz;
    const oreal d = tr * (oreal)(1.0 / 3.0);
    const oreal U3 = (oreal)3 * U;
    for (int i = 0; i < L->Q; i++) {
        const int ex = L->e[i][0], ey = L->e[i][1], ez = L->e[i][2];
        oreal q = 0.0, o = 0.0;
        if (ex != 0) q = q + Pxx;
        if (ey != 0) q = q + Pyy;
        if (ez != 0) q = q + Pzz;
        q = q - d;
        if (ex * ey > 0) o = o + Mxy; else if (ex * ey < 0) o = o - Mxy;
        if (ex * ez > 0) o = o + Mxz; else if (ex * ez < 0) o = o - Mxz;
        if (ey * ez > 0) o = o + Myz; else if (ey * ez < 0) o = o - Myz;
        oreal t = S;
        if (ex > 0) t = S + U3; else if (ex < 0) t = S - U3;
        g[i] = (oreal)L->t[i] * (t + ((oreal)4.5 * q + (oreal)9 * o));
    }
}

void orc_plain_step(orc_plain *s, int nsteps)
{
    const lattice *L = &s->L;
    const int Q = L->Q, D = L->D, nx = s->nx, ny = s->ny, nz = s->nz;
    const size_t N = s->N, rows = (size_t)ny * nz;
    for (int it = 0; it < nsteps; it++) {
        s->step++;
        memset(s->rowF, 0, rows * 3 * sizeof(double));
        /* 1. push: every source population either streams to its neighbour or bounces */
#pragma omp parallel for
        for (long r = 0; r < (long)rows; r++) {
            const int y = (int)(r % ny), z = (int)(r / ny);
            double *F = s->rowF + 3 * r;
            for (int x = 0; x < nx; x++) {
                const size_t a = (size_t)x + (size_t)nx * r;
                if (s->solid[a]) continue;
                for (int i = 0; i < Q; i++) {
                    const int *e = L->e[i];
                    const oreal v = s->h[i * N + a];
                    int dx = x + e[0], dy = y + e[1], dz = z + e[2];
                    int in = 1;
                    if (s->xper) dx = wrapi(dx, nx); else if (dx < 0 || dx >= nx) in = 0;
                    if (s->yper) dy = wrapi(dy, ny); else if (dy < 0 || dy >= ny) in = 0;
                    if (D == 3) { if (s->zper) dz = wrapi(dz, nz); else if (dz < 0 || dz >= nz) in = 0; }
                    size_t d = in ? (size_t)dx + (size_t)nx * ((size_t)dy + (size_t)ny * dz) : 0;
                    if (in && !s->solid[d]) {
                        s->g[i * N + d] = v;
                        continue;
                    }
                    size_t dummy;
                    int c = classify(s, x, y, z, L->bb[i], &dummy);
                    if (c == C_WALL || c == C_MASK) {
                        s->g[L->bb[i] * N + a] = v;
                        if (c == C_MASK) {
                            F[0] += 2.0 * v * (double)e[0];
                            F[1] += 2.0 * v * (double)e[1];
                            F[2] += 2.0 * v * (double)e[2];
                        }
                    }
                }
            }
        }
        /* 2. outlet (zero-gradient) and regularised Zou-He inlet */
        if (!s->xper) {
#pragma omp parallel for
            for (long r = 0; r < (long)rows; r++) {
                const int y = (int)(r % ny), z = (int)(r / ny);
                size_t d = (size_t)(nx - 1) + (size_t)nx * r;
                if (!s->solid[d])
                    for (int i = 0; i < Q; i++) {
                        size_t src;
                        if (classify(s, nx - 1, y, z, i, &src) == C_OUTLET) s->g[i * N + d] = s->h[i * N + src];
                    }
                d = (size_t)nx * r;
                if (!s->solid[d]) {
                    oreal gv[27];
                    for (int i = 0; i < Q; i++) gv[i] = s->g[i * N + d];
                    zou_he(L, gv, (oreal)s->uin[r]);
                    regularize_inlet(L, gv, (oreal)s->uin[r]);
                    for (int i = 0; i < Q; i++) s->g[i * N + d] = gv[i];
                }
            }
        }
        /* 3. moments (reference term order), partial bounce-back force, BGK collision */
#pragma omp parallel for
        for (long r = 0; r < (long)rows; r++) {
            double *F = s->rowF + 3 * r;
            for (int x = 0; x < nx; x++) {
                const size_t a = (size_t)x + (size_t)nx * r;
                if (s->solid[a]) {
                    for (int i = 0; i < Q; i++) s->h[i * N + a] = 0.0;
                    s->mp[a] = s->mu[a] = s->mv[a] = s->mw[a] = 0.0;
                    continue;
                }
                oreal gv[27];
                oreal p = 0.0, u = 0.0, v = 0.0, w = 0.0;
                for (int i = 0; i < Q; i++) {
                    gv[i] = s->g[i * N + a];
                    p = p + gv[i] / (oreal)3;
                    const int *e = L->e[i];
                    if (e[0] > 0) u = u + gv[i]; else if (e[0] < 0) u = u - gv[i];
                    if (e[1] > 0) v = v + gv[i]; else if (e[1] < 0) v = v - gv[i];
                    if (e[2] > 0) w = w + gv[i]; else if (e[2] < 0) w = w - gv[i];
                }
                s->mp[a] = p; s->mu[a] = u; s->mv[a] = v; s->mw[a] = w;
                const oreal sf = (oreal)(s->s ? s->s[a] : 0.0);
                if (s->s) {
                    F[0] += 2.0 * sf * u;
                    F[1] += 2.0 * sf * v;
                    F[2] += 2.0 * sf * w;
                }
                for (int i = 0; i < Q; i++) {
                    oreal feq = feq_p(L, i, p, u, v, w);
                    oreal c = gv[i] - (gv[i] - feq) / (oreal)(s->tau0 + 0.5);
                    if (s->s) c = ((oreal)1 - sf) * c + sf * gv[L->bb[i]];
                    s->h[i * N + a] = c;
                }
            }
        }
        if (s->force_interval > 0 && s->step % s->force_interval == 0) {
            double F[3] = {0.0, 0.0, 0.0};
            for (size_t r = 0; r < rows; r++) {
                F[0] += s->rowF[3 * r]; F[1] += s->rowF[3 * r + 1]; F[2] += s->rowF[3 * r + 2];
            }
            push_force(&s->frec, &s->nrec, &s->caprec, s->step, F);
        }
    }
}

void orc_plain_fields(const orc_plain *s, double *p, double *u, double *v, double *w)
{
    size_t b = s->N * sizeof(double);
    if (p) memcpy(p, s->mp, b);
    if (u) memcpy(u, s->mu, b);
    if (v) memcpy(v, s->mv, b);
    if (w) memcpy(w, s->mw, b);
}
void orc_plain_populations(const orc_plain *s, double *out)
{
    for (size_t k = 0; k < s->N * s->L.Q; k++) out[k] = s->h[k];
}
void orc_plain_set_populations(orc_plain *s, const double *in)
{
    for (size_t k = 0; k < s->N * s->L.Q; k++) s->h[k] = (oreal)in[k];
}
void orc_plain_last_g(const orc_plain *s, double *out)
{
    for (size_t k = 0; k < s->N * s->L.Q; k++) out[k] = s->g[k];
}
void orc_plain_inlet(const orc_plain *s, double *uin) { memcpy(uin, s->uin, (size_t)s->ny * s->nz * sizeof(double)); }
int orc_plain_forces(const orc_plain *s, double *out, int max)
{
    int n = s->nrec < max ? s->nrec : max;
    if (out) memcpy(out, s->frec, (size_t)n * 4 * sizeof(double));
    return s->nrec;
}

#ifndef ORC_FLOAT
/* ==================================================================================== */
/* Discrete adjoint of the plain-mode step (FP64).  No reference counterpart.            */
/*                                                                                      */
/* Step m (1-based) maps H_{m-1} (post-collision) to G_m (streamed + boundaries) and     */
/* H_m = (1-s) C(G_m) + s G_m[bb], with drag Fx_m = link part (2 H_{m-1} e_x on links    */
/* into the mask) + partial-bounce-back part (2 s u_x(G_m)).  J = scale * mean_{m in    */
/* [n0,n1]} Fx_m.  The reverse sweep carries mu_m = dJ/dG_m; lambda_m = dJ/dH_m is the */
/* transpose of streaming applied to mu_{m+1} (scatter form, serial).                    */
/* ==================================================================================== */
static double window_weight(int m, int n0, int n1, double scale)
{
    return (m >= n0 && m <= n1) ? scale / (double)(n1 - n0 + 1) : 0.0;
}

/* lam = S^T mu (+ link-force partials of the same step, weight alpha) */
static void adj_stream_T(const orc_plain *s, const double *mu, double alpha, double *lam)
{
    const lattice *L = &s->L;
    const int Q = L->Q, nx = s->nx, ny = s->ny, nz = s->nz;
    const size_t N = s->N;
    memset(lam, 0, N * Q * sizeof(double));
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const size_t d = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                if (s->solid[d]) continue;
                for (int j = 0; j < Q; j++) {
                    size_t src = 0;
                    const double m = mu[j * N + d];
                    switch (classify(s, x, y, z, j, &src)) {
                    case C_NORMAL:
                    case C_OUTLET: lam[j * N + src] += m; break;
                    case C_WALL: lam[L->bb[j] * N + d] += m; break;
                    case C_MASK:
                        lam[L->bb[j] * N + d] += m + alpha * 2.0 * (double)L->e[L->bb[j]][0];
                        break;
                    default: break; /* Zou-He entries read no population of H */
                    }
                }
            }
}

/* mu = (collision + partial bounce-back + Zou-He)^T lam at every node, for step m whose
 * pre-collision populations are G; accumulates dJ/ds and dJ/du_in. */
static void adj_local(const orc_plain *s, const double *G, const double *lam, double alpha,
                      double *mu, double *dJ_ds, double *dJ_duin)
{
    const lattice *L = &s->L;
    const int Q = L->Q, nx = s->nx;
    const size_t N = s->N;
    const double om = 1.0 / (s->tau0 + 0.5);
    for (size_t a = 0; a < N; a++) {
        if (s->solid[a]) {
            for (int j = 0; j < Q; j++) mu[j * N + a] = 0.0;
            continue;
        }
        double g[27], l[27], m[27];
        double p = 0.0, u[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < Q; i++) {
            g[i] = G[i * N + a];
            l[i] = lam[i * N + a];
            p += g[i] / 3.0;
            for (int d = 0; d < 3; d++) u[d] += g[i] * (double)L->e[i][d];
        }
        const double sf = s->s ? s->s[a] : 0.0;
        /* dJ/ds: H_i = (1-s) c_i + s g_bb(i) ; Fx = 2 s u_x */
        if (s->s) {
            double ds = alpha * 2.0 * u[0];
            for (int i = 0; i < Q; i++) {
                double feq = feq_ref(L, i, p, u[0], u[1], u[2]);
                double c = g[i] - (g[i] - feq) * om;
                ds += l[i] * (g[L->bb[i]] - c);
            }
            dJ_ds[a] += ds;
        }
        /* collision transpose applied to (1-s) lam */
        double Lp = 0.0, Lu[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < Q; i++) {
            const double li = (1.0 - sf) * l[i], w = L->t[i];
            double eu = 0.0;
            for (int d = 0; d < 3; d++) eu += (double)L->e[i][d] * u[d];
            Lp += li * w;
            for (int d = 0; d < 3; d++) Lu[d] += li * w * (3.0 * L->e[i][d] + 9.0 * eu * L->e[i][d] - 3.0 * u[d]);
        }
        for (int j = 0; j < Q; j++) {
            double eL = 0.0;
            for (int d = 0; d < 3; d++) eL += (double)L->e[j][d] * Lu[d];
            m[j] = (1.0 - om) * (1.0 - sf) * l[j] + om * (Lp + eL) + sf * l[L->bb[j]] + alpha * 2.0 * sf * L->e[j][0];
        }
        /* inlet nodes: transpose of the regularisation g -> A g + 3 w e_x U with the explicit
         * matrix A_ij = w_i (1 + 4.5 Q_i:Q_j), then the Zou-He transpose */
        const int x = (int)(a % nx);
        if (!s->xper && x == 0) {
            const size_t row = a / nx;
            const int D = L->D;
            double mr[27];
            for (int j = 0; j < Q; j++) {
                double acc = 0.0;
                for (int i = 0; i < Q; i++) {
                    double qq = 0.0;
                    for (int p1 = 0; p1 < D; p1++)
                        for (int p2 = 0; p2 < D; p2++) {
                            const double qi = (double)(L->e[i][p1] * L->e[i][p2]) - (p1 == p2 ? 1.0 / 3.0 : 0.0);
                            const double qj = (double)(L->e[j][p1] * L->e[j][p2]) - (p1 == p2 ? 1.0 / 3.0 : 0.0);
                            qq += qi * qj;
                        }
                    acc += m[i] * L->t[i] * (1.0 + 4.5 * qq);
                }
                mr[j] = acc;
            }
            for (int i = 0; i < Q; i++) dJ_duin[row] += m[i] * 3.0 * L->t[i] * (double)L->e[i][0];
            for (int j = 0; j < Q; j++) m[j] = mr[j];
            for (int i = 0; i < Q; i++) {
                if (L->e[i][0] <= 0) continue;
                const double mi = m[i];
                m[L->bb[i]] += mi;
                for (int k = 0; k < Q; k++)
                    if (L->e[k][0] == 0)
                        m[k] += mi * (-18.0 * L->t[i]) * ((double)(L->e[i][1] * L->e[k][1] + L->e[i][2] * L->e[k][2]));
                dJ_duin[row] += mi * 6.0 * L->t[i];
                m[i] = 0.0;
            }
        }
        for (int j = 0; j < Q; j++) mu[j * N + a] = m[j];
    }
}

static double run_objective(orc_plain *s, int n1, int n0, double scale)
{
    orc_plain_step(s, n1);
    double J = 0.0;
    for (int r = 0; r < s->nrec; r++) {
        int m = (int)s->frec[4 * r];
        J += window_weight(m, n0, n1, scale) * s->frec[4 * r + 1];
    }
    return J;
}

double orc_objective(const orc_cfg *cfg, const unsigned char *mask, const double *sfrac,
                     const double *xi, const double *uin, int nsteps, int n0, int n1, double cd_scale)
{
    orc_cfg c = *cfg;
    c.force_interval = 1;
    orc_plain *s = orc_plain_create(&c, mask, sfrac, xi, uin);
    if (!s) return NAN;
    (void)nsteps;
    double J = run_objective(s, n1, n0, cd_scale);
    orc_plain_destroy(s);
    return J;
}

double orc_adjoint(const orc_cfg *cfg, const unsigned char *mask, const double *sfrac,
                   const double *xi, const double *uin, int nsteps, int n0, int n1,
                   double cd_scale, int K, double *dJ_ds, double *dJ_duin)
{
    orc_cfg c = *cfg;
    c.force_interval = 1;
    (void)nsteps;
    if (K <= 0) K = n1;
    orc_plain *s = orc_plain_create(&c, mask, sfrac, xi, uin);
    if (!s) return NAN;
    const int Q = s->L.Q;
    const size_t N = s->N, SZ = N * Q, rows = (size_t)s->ny * s->nz;
    const int nseg = (n1 + K - 1) / K;
    /* forward with checkpoints H_{kK} */
    double *ckpt = (double *)malloc(SZ * nseg * sizeof(double));
    for (int k = 0; k < nseg; k++) {
        memcpy(ckpt + SZ * k, s->h, SZ * sizeof(double));
        int n = (k + 1) * K <= n1 ? K : n1 - k * K;
        orc_plain_step(s, n);
    }
    double J = 0.0;
    for (int r = 0; r < s->nrec; r++)
        J += window_weight((int)s->frec[4 * r], n0, n1, cd_scale) * s->frec[4 * r + 1];
    memset(dJ_ds, 0, N * sizeof(double));
    memset(dJ_duin, 0, rows * sizeof(double));
    double *Gseg = (double *)malloc(SZ * K * sizeof(double));
    double *lam = (double *)malloc(SZ * sizeof(double));
    double *mu_next = (double *)calloc(SZ, sizeof(double));
    double *mu = (double *)malloc(SZ * sizeof(double));
    const int fi = s->force_interval;
    s->force_interval = 0;
    for (int k = nseg - 1; k >= 0; k--) {
        const int m0 = k * K;                 /* checkpoint holds H_{m0} */
        const int n = (m0 + K <= n1) ? K : n1 - m0;
        memcpy(s->h, ckpt + SZ * k, SZ * sizeof(double));
        for (int t = 0; t < n; t++) {         /* recompute G_{m0+1..m0+n} */
            orc_plain_step(s, 1);
            memcpy(Gseg + SZ * t, s->g, SZ * sizeof(double));
        }
        for (int t = n - 1; t >= 0; t--) {
            const int m = m0 + t + 1;
            adj_stream_T(s, mu_next, window_weight(m + 1, n0, n1, cd_scale), lam);
            adj_local(s, Gseg + SZ * t, lam, window_weight(m, n0, n1, cd_scale), mu, dJ_ds, dJ_duin);
            double *tmp = mu_next; mu_next = mu; mu = tmp;
        }
    }
    s->force_interval = fi;
    /* lambda_0 and the initial condition H_0 = feq(0, u_in + perturbation) */
    adj_stream_T(s, mu_next, window_weight(1, n0, n1, cd_scale), lam);
    const lattice *L = &s->L;
    for (size_t a = 0; a < N; a++) {
        if (s->solid[a]) continue;
        const size_t row = a / s->nx;
        double u = s->uin[row] + c.perturb * (xi ? xi[a] : orc_xi(c.seed, (uint64_t)a));
        double acc = 0.0;
        for (int i = 0; i < Q; i++) {
            double ex = (double)L->e[i][0];
            acc += lam[i * N + a] * L->t[i] * (3.0 * ex + 9.0 * ex * u * ex - 3.0 * u);
        }
        dJ_duin[row] += acc;
    }
    free(ckpt); free(Gseg); free(lam); free(mu_next); free(mu);
    orc_plain_destroy(s);
    return J;
}
#endif /* ORC_FLOAT */
