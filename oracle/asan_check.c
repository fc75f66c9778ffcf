/*
 * asan_check.c -- drives every entry point of the CPU oracle on small cases, for the
 * AddressSanitizer + UndefinedBehaviorSanitizer build (make asan_check).  TEST
 * INFRASTRUCTURE ONLY: run by tests/test_oracle_sanitizer.py; never part of the product.
 * Exits 0 when every case ran; a sanitizer report aborts the program with a non-zero status.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

static orc_cfg cfg(int Q, int nx, int ny, int nz)
{
    orc_cfg c;
    memset(&c, 0, sizeof(c));
    c.Q = Q; c.nx = nx; c.ny = ny; c.nz = Q == 27 ? nz : 1;
    c.tau0 = 0.06; c.u0 = 0.05;
    c.body = ORC_BODY_SPHERE;
    c.cx = nx / 3.0; c.cy = (ny - 1) / 2.0; c.cz = (c.nz - 1) / 2.0; c.r2 = 4.0;
    c.seed = 5374857; c.perturb = 1e-3; c.force_interval = 2;
    return c;
}

static double checksum(const double *a, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += a[i];
    return s;
}

static int reference_cases(void)
{
    const int shapes[3][4] = {{9, 20, 12, 1}, {27, 12, 9, 8}, {9, 18, 10, 1}};
    for (int c = 0; c < 3; c++) {
        orc_cfg k = cfg(shapes[c][0], shapes[c][1], shapes[c][2], shapes[c][3]);
        k.variant = c == 2;
        orc_ref *s = orc_ref_create(&k, NULL);
        if (!s) return 1;
        orc_ref_step(s, 3);
        for (int st = 0; st < 4; st++) orc_ref_stage(s, st);
        int lx, ly, lz, g;
        orc_ref_extents(s, &lx, &ly, &lz, &g);
        const size_t N = (size_t)k.nx * k.ny * k.nz, P = (size_t)lx * ly * lz * k.Q;
        double *f = malloc(sizeof(double) * k.Q * N), *fl = malloc(sizeof(double) * 4 * N);
        double *raw = malloc(sizeof(double) * P), F[3], rec[40];
        orc_ref_fields(s, fl, fl + N, fl + 2 * N, fl + 3 * N);
        orc_ref_populations(s, f);
        memcpy(raw, orc_ref_raw(s), sizeof(double) * P);
        orc_ref_set_raw(s, raw);
        orc_ref_collision(s);
        orc_ref_boundary_fills(s);
        orc_ref_usrbc(s);
        orc_ref_mem(s, F);
        orc_ref_stream(s);
        orc_ref_update(s);
        const int n = orc_ref_forces(s, rec, 10);
        printf("ref Q%d variant %d: %d force records, checksum %.6e\n", k.Q, k.variant, n, checksum(f, k.Q * N) + F[0]);
        free(f); free(fl); free(raw);
        orc_ref_destroy(s);
    }
    return 0;
}

static int plain_cases(void)
{
    for (int c = 0; c < 4; c++) {
        const int Q = c % 2 ? 27 : 9;
        orc_cfg k = cfg(Q, 14, 10, 8);
        const size_t N = (size_t)k.nx * k.ny * k.nz;
        unsigned char *mask = NULL;
        double *sfrac = NULL, *xi = NULL, *uin = NULL;
        if (c >= 2) {   /* mask body, solid fraction, injected perturbation and inlet profile, periodic y/z */
            k.body = ORC_BODY_MASK;
            k.inlet = ORC_INLET_PROFILE;
            k.yper = 1; k.zper = Q == 27;
            mask = calloc(N, 1);
            sfrac = calloc(N, sizeof(double));
            xi = calloc(N, sizeof(double));
            uin = calloc((size_t)k.ny * k.nz, sizeof(double));
            for (size_t n = 0; n < N; n++) {
                mask[n] = n % 37 == 5;
                sfrac[n] = n % 11 == 3 ? 0.4 : 0.0;
                xi[n] = (double)(n % 7) / 7.0 - 0.5;
            }
            for (int r = 0; r < k.ny * k.nz; r++) uin[r] = 0.04 + 1e-3 * (r % 3);
        } else {
            k.inlet = ORC_INLET_LEGENDRE;
            k.ncoef = 3; k.coef[0] = 0.05; k.coef[1] = 0.01; k.coef[2] = -0.005;
        }
        orc_plain *s = orc_plain_create(&k, mask, sfrac, xi, uin);
        if (!s) return 1;
        orc_plain_step(s, 4);
        double *f = malloc(sizeof(double) * Q * N), *fl = malloc(sizeof(double) * 4 * N);
        double *in = malloc(sizeof(double) * k.ny * k.nz), rec[40];
        orc_plain_populations(s, f);
        orc_plain_set_populations(s, f);
        orc_plain_step(s, 1);
        orc_plain_last_g(s, f);
        orc_plain_fields(s, fl, fl + N, fl + 2 * N, fl + 3 * N);
        orc_plain_inlet(s, in);
        const int n = orc_plain_forces(s, rec, 10);
        printf("plain case %d Q%d: %d force records, checksum %.6e\n", c, Q, n, checksum(f, Q * N));
        orc_plain_destroy(s);
        if (c >= 2) {
            double *ds = calloc(N, sizeof(double)), *du = calloc((size_t)k.ny * k.nz, sizeof(double));
            const double J = orc_adjoint(&k, mask, sfrac, xi, uin, 8, 3, 8, 1.0, 3, ds, du);
            const double J2 = orc_objective(&k, mask, sfrac, xi, uin, 8, 3, 8, 1.0);
            printf("adjoint Q%d: J %.6e (forward %.6e), dJ/ds checksum %.6e\n", Q, J, J2, checksum(ds, N));
            free(ds); free(du);
        }
        free(f); free(fl); free(in); free(mask); free(sfrac); free(xi); free(uin);
    }
    return 0;
}

int main(void)
{
    double fe[27];
    int e[81], bb[27];
    double w[27];
    for (int Q = 9; Q <= 27; Q += 18) {
        if (orc_lattice(Q, e, w, bb) || orc_feq(Q, 0.01, 0.05, 0.01, 0.0, fe)) return 1;
    }
    printf("rng %llu %.6f %d\n", (unsigned long long)orc_hash(1, 2), orc_xi(3, 4), orc_ran100(5, 6));
    if (reference_cases() || plain_cases()) return 1;
    printf("asan_check: ok\n");
    return 0;
}
