// adjoint.cpp -- checkpointed discrete-adjoint driver (lbm_adjoint; no reference counterpart).
//
// Forward: steps 1..n1 from the initial state, drag recorded every step, the state saved
// every K steps (checkpoints H_0, H_K, ...).  Reverse: for each segment, newest first, the
// checkpoint is restored and the segment's states H_{m0} .. H_{m0+n-1} are recomputed into
// a device-resident store (up to K states: with 288 GB of HBM a whole segment of a 1.8 M-node
// D3Q27 FP32 run, ~48 GB at K = 250, stays on the GPU), then one reverse kernel per step
// turns mu_{m+1} into mu_m.  Cost: the recompute is one forward pass; each reverse step
// moves 3Q values per node against the forward's 2Q.
#include <cmath>
#include <cstring>

#include "internal.hpp"

using namespace lbm;

static int aerr(lbm_ctx *c, int code, const std::string &m) {
    c->err = m;
    return code;
}

#define ATRY(expr)                                                                                     \
    do {                                                                                               \
        hipError_t e__ = (expr);                                                                       \
        if (e__ != hipSuccess) {                                                                       \
            rc = aerr(c, LBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e__));             \
            goto done;                                                                                 \
        }                                                                                              \
    } while (0)

static double weight(int m, int n0, int n1, double scale) {
    return (m >= n0 && m <= n1) ? scale / (double)(n1 - n0 + 1) : 0.0;
}

FusedArgs ctx_fused_args(const lbm_ctx *c);
StepBufs step_bufs(const lbm_ctx *c, const void *A, void *B);

extern "C" int lbm_adjoint(lbm_ctx *c, int n0, int n1, double scale, int K, double *J, double *dJ_ds,
                           double *dJ_duin, double *ms_forward, double *ms_adjoint) {
    if (!c) return LBM_ERR_INVALID;
    if (c->cfg.bc_mode != LBM_BC_PLAIN) return aerr(c, LBM_ERR_INVALID, "the adjoint covers the plain boundary mode");
    if (c->prec != LBM_FP64 && c->prec != LBM_FP32) return aerr(c, LBM_ERR_INVALID, "the adjoint needs FP64 or FP32 storage");
    if (!c->initialized || c->step != 0) return aerr(c, LBM_ERR_STATE, "lbm_adjoint needs a freshly initialised context (step 0)");
    if (n0 < 1 || n1 < n0) return aerr(c, LBM_ERR_INVALID, "need 1 <= n0 <= n1");
    if (K <= 0 || K > n1) K = n1;
    const Geo &g = c->geo;
    const int Q = c->Q;
    const size_t esz = c->prec == LBM_FP64 ? 8 : 4;
    const size_t SZ = esz * Q * g.plane;
    const int nseg = (n1 + K - 1) / K;
    int rc = LBM_OK;
    void *ckpt = nullptr, *store = nullptr, *mu0 = nullptr, *mu1 = nullptr;
    double *dJds = nullptr, *dJdu = nullptr, *tmp = nullptr, *dxi = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    float tf = 0.f, ta = 0.f;
    const FusedArgs fa = ctx_fused_args(c);
    const int fi_saved = c->cfg.force_interval;
    std::vector<double> hist;
    size_t freeb = 0, totalb = 0;

    (void)hipMemGetInfo(&freeb, &totalb);
    if ((double)SZ * (nseg + K + 2) + 8.0 * g.plane * 2 > 0.95 * (double)freeb)
        return aerr(c, LBM_ERR_NOMEM, "adjoint store (" + std::to_string((double)SZ * (nseg + K + 2) / 1e9) +
                                         " GB) exceeds free device memory; lower checkpoint_every");
    ATRY(hipMalloc(&ckpt, SZ * nseg));
    ATRY(hipMalloc(&store, SZ * K));
    ATRY(hipMalloc(&mu0, SZ));
    ATRY(hipMalloc(&mu1, SZ));
    ATRY(hipMalloc(&dJds, sizeof(double) * g.plane));
    ATRY(hipMalloc(&dJdu, sizeof(double) * g.ny * g.nz));
    ATRY(hipMemsetAsync(mu0, 0, SZ, c->stream));
    ATRY(hipMemsetAsync(mu1, 0, SZ, c->stream));
    ATRY(hipMemsetAsync(dJds, 0, sizeof(double) * g.plane, c->stream));
    ATRY(hipMemsetAsync(dJdu, 0, sizeof(double) * g.ny * g.nz, c->stream));
    if (!c->xi.empty()) {
        ATRY(hipMalloc(&dxi, sizeof(double) * c->N));
        ATRY(hipMemcpyAsync(dxi, c->xi.data(), sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
    }
    ATRY(hipEventCreate(&e0));
    ATRY(hipEventCreate(&e1));
    ATRY(hipEventCreate(&e2));

    // ---- forward with checkpoints, drag every step ----
    c->cfg.force_interval = 1;
    ATRY(hipEventRecord(e0, c->stream));
    for (int k = 0; k < nseg; k++) {
        ATRY(hipMemcpyAsync((char *)ckpt + SZ * k, c->A, SZ, hipMemcpyDeviceToDevice, c->stream));
        const int n = (k + 1) * K <= n1 ? K : n1 - k * K;
        rc = lbm_step(c, n);
        if (rc != LBM_OK) goto done;
    }
    ATRY(hipEventRecord(e1, c->stream));

    // ---- reverse sweep ----
    {
        void *mun = mu0, *mu = mu1;   // mun holds mu_{m+1} (zero beyond n1)
        for (int k = nseg - 1; k >= 0; k--) {
            const int m0 = k * K;
            const int n = (m0 + K <= n1) ? K : n1 - m0;
            ATRY(hipMemcpyAsync(store, (char *)ckpt + SZ * k, SZ, hipMemcpyDeviceToDevice, c->stream));
            for (int t = 0; t + 1 < n; t++)
                ATRY(plain_step(Q, c->prec, c->V, step_bufs(c, (char *)store + SZ * t, (char *)store + SZ * (t + 1)),
                                false, false, fa, c->stream));
            for (int t = n - 1; t >= 0; t--) {
                const int m = m0 + t + 1;
                AdjArgs aa;
                aa.f = fa;
                aa.alpha = weight(m, n0, n1, scale);
                aa.alpha_next = weight(m + 1, n0, n1, scale);
                ATRY(adj_step(Q, c->prec, (char *)store + SZ * t, mun, mu, c->flags, c->dsfrac, c->duin, dJds, dJdu, c->rowclean, c->adj_all_fast, aa,
                              c->stream));
                std::swap(mun, mu);
            }
        }
        ATRY(adj_init(Q, c->prec, mun, c->flags, c->duin, dxi, c->cfg.perturb, c->cfg.seed, weight(1, n0, n1, scale),
                      dJdu, fa, c->stream));
    }
    ATRY(hipEventRecord(e2, c->stream));
    ATRY(hipEventSynchronize(e2));
    ATRY(hipEventElapsedTime(&tf, e0, e1));
    ATRY(hipEventElapsedTime(&ta, e1, e2));
    if (ms_forward) *ms_forward = tf;
    if (ms_adjoint) *ms_adjoint = ta;

    // ---- outputs ----
    if (J) {
        hist.resize(4 * (size_t)c->nrec);
        ATRY(hipMemcpy(hist.data(), c->hist, sizeof(double) * hist.size(), hipMemcpyDeviceToHost));
        double acc = 0.0;
        for (int r = 0; r < c->nrec; r++) acc += weight((int)hist[4 * r], n0, n1, scale) * hist[4 * r + 1];
        *J = acc;
    }
    if (dJ_ds) {
        ATRY(hipMalloc(&tmp, sizeof(double) * c->N));
        ATRY(adj_reduce_ds(c->prec, dJds, tmp, g, c->stream));
        ATRY(hipMemcpyAsync(dJ_ds, tmp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
    }
    if (dJ_duin) ATRY(hipMemcpyAsync(dJ_duin, dJdu, sizeof(double) * g.ny * g.nz, hipMemcpyDeviceToHost, c->stream));
    ATRY(hipStreamSynchronize(c->stream));

done:
    c->cfg.force_interval = fi_saved;
    (void)hipStreamSynchronize(c->stream);
    for (void *p : {ckpt, store, mu0, mu1, (void *)dJds, (void *)dJdu, (void *)tmp, (void *)dxi})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : {e0, e1, e2})
        if (e) (void)hipEventDestroy(e);
    return rc;
}
