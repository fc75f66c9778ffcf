// lbm_capi.cpp -- implementation of include/lbm_capi.h (host side; kernels live in *.hip).
//
// The context owns every device allocation and one non-blocking HIP stream (default
// priority).  The reference's set-up (System::System + LBM::LBM + LBM::Initialize,
// 3D MPI+KOKKOS+MPIIO/System.cpp:5-23, lbm.hpp:146-219, lbm.cpp:36-256) becomes
// lbm_create + lbm_initialize; its time loop body becomes lbm_step.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "internal.hpp"
#include "lattice.hpp"

using namespace lbm;

static thread_local std::string t_create_err;

static bool ref_mode(int bc) { return bc == LBM_BC_REFERENCE || bc == LBM_BC_REFERENCE_SN; }

static int set_err(lbm_ctx *c, int code, const std::string &m) {
    if (c) c->err = m;
    else t_create_err = m;
    return code;
}

#define HIP_TRY(c, expr)                                                                               \
    do {                                                                                               \
        hipError_t e__ = (expr);                                                                       \
        if (e__ != hipSuccess)                                                                         \
            return set_err((c), LBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e__));     \
    } while (0)

template <int Q>
static void table(int *e, double *w, int *opp) {
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        if (e) {
            e[3 * i] = Lat<Q>::ex(i);
            e[3 * i + 1] = Lat<Q>::ey(i);
            e[3 * i + 2] = Lat<Q>::ez(i);
        }
        if (w) w[i] = Lat<Q>::w(i);
        if (opp) opp[i] = Lat<Q>::opp(i);
    });
}

extern "C" int lbm_lattice_table(int lattice, int *e, double *w, int *opp) {
    if (lattice == 9) table<9>(e, w, opp);
    else if (lattice == 27) table<27>(e, w, opp);
    else return LBM_ERR_INVALID;
    return LBM_OK;
}

extern "C" int lbm_config_default(lbm_config *c, int lattice) {
    if (!c || (lattice != 9 && lattice != 27)) return LBM_ERR_INVALID;
    std::memset(c, 0, sizeof(*c));
    const double cs2 = 1.0 / 3.0;
    c->lattice = lattice;
    c->rho0 = 1.0;
    c->precision = LBM_FP64;
    c->bc_mode = LBM_BC_REFERENCE;
    c->body = LBM_BODY_SPHERE;
    c->inlet = LBM_INLET_UNIFORM;
    c->seed = 5374857ull;   // the reference's Random_XorShift64_Pool seed
    if (lattice == 27) {
        // input.in: rho R Re / u0 Time inter / nx ny nz = 1.0 8.0 100 / 0.1 .. / 512 128 128
        const double R = 8.0, Re = 100.0;
        c->nx = 512; c->ny = 128; c->nz = 128;
        c->u0 = 0.1;
        c->tau0 = c->u0 * R / Re / cs2;                       // System.cpp:19
        c->cx = c->cy = c->cz = (double)((c->nx - 1) / 8);    // lbm.cpp:238 (integer division)
        const double r = (c->nx - 1) / 64.0;
        c->r2 = r * r;
        c->force_interval = 0;                                // the 3D reference logs no force
    } else {
        // MPI Kokkos/System.cpp:9-17 and main.cpp:8-9: nx 512, ny 128, R = int(0.1*ny)
        c->nx = 512; c->ny = 128; c->nz = 1;
        c->u0 = 0.1;
        const int R = (int)(0.1 * c->ny);
        c->tau0 = c->u0 * R / 100.0 / cs2;
        c->cx = 0.5 * (c->nx - 1);                            // lbm.cpp:108
        c->cy = 0.5 * (c->ny - 1);
        const double r = 0.03125 * (c->nx - 1) / 2.0;
        c->r2 = r * r;
        c->force_interval = 1;                                // test.dat every step, lbm.cpp:398-407
    }
    return LBM_OK;
}

extern "C" const char *lbm_last_error(const lbm_ctx *ctx) { return ctx ? ctx->err.c_str() : t_create_err.c_str(); }

static void free_all(lbm_ctx *c) {
    void *ptrs[] = {c->A, c->B, c->flags, c->dsfrac, c->duin, c->macros, c->block_slot, c->partials, c->hist, c->fix,
                    c->fixlk, c->rowwin, c->wlk, c->link, c->linklk, c->link_slot, c->rowclean, c->rf, c->rft, c->rmac, c->rprof, c->rusr, c->rfpre, c->rsurf,
                    c->snap_dev, c->dbad, c->tb_partials, c->tb_lk};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->A = c->B = nullptr;
    c->flags = nullptr;
    c->dsfrac = nullptr;
    c->duin = c->macros = c->partials = c->hist = nullptr;
    c->block_slot = nullptr;
    c->fix = c->fixlk = nullptr;
    c->rowwin = nullptr;
    c->wlk = nullptr;
    c->link = c->linklk = nullptr;
    c->link_slot = nullptr;
    c->rowclean = nullptr;
    c->nfix = c->nlink = 0;
    c->snap_dev = nullptr;
    c->dbad = nullptr;
    c->tb_partials = nullptr;
    c->tb_lk = nullptr;
    c->tb_k = 0;
    c->rf = c->rft = c->rmac = c->rprof = c->rfpre = nullptr;
    c->rsurf = nullptr;
    c->nsurf = 0;
    c->rusr = nullptr;
    if (c->snap_host) (void)hipHostFree(c->snap_host);
    c->snap_host = nullptr;
    c->snap_elems = 0;
}

extern "C" int lbm_create(const lbm_config *cfg, lbm_ctx **out) {
    if (!cfg || !out) return set_err(nullptr, LBM_ERR_INVALID, "null argument");
    *out = nullptr;
    const int Q = cfg->lattice;
    if (Q != 9 && Q != 27) return set_err(nullptr, LBM_ERR_INVALID, "lattice must be 9 (D2Q9) or 27 (D3Q27)");
    const int nz = Q == 9 ? 1 : cfg->nz;
    if (cfg->nx < 3 || cfg->ny < 3 || (Q == 27 && nz < 3) || (Q == 9 && cfg->nz > 1))
        return set_err(nullptr, LBM_ERR_INVALID, "extents: need nx, ny (and nz for D3Q27) >= 3, nz = 1 for D2Q9");
    if (!(cfg->tau0 > 0.0) || !std::isfinite(cfg->tau0))
        return set_err(nullptr, LBM_ERR_INVALID, "tau0 must be > 0 (relaxation time tau0 + 0.5 > 0.5)");
    if (cfg->precision < LBM_FP64 || cfg->precision > LBM_BF16S)
        return set_err(nullptr, LBM_ERR_INVALID, "unknown precision");
    if (!ref_mode(cfg->bc_mode) && cfg->bc_mode != LBM_BC_PLAIN)
        return set_err(nullptr, LBM_ERR_INVALID, "unknown bc_mode");
    if (cfg->bc_mode == LBM_BC_REFERENCE_SN && Q != 9)
        return set_err(nullptr, LBM_ERR_INVALID, "LBM_BC_REFERENCE_SN is the single-node 2D variant: lattice 9 only");
    if (ref_mode(cfg->bc_mode) && cfg->precision != LBM_FP64)
        return set_err(nullptr, LBM_ERR_INVALID, "reference boundary mode runs in FP64 only");
    if (cfg->body < LBM_BODY_NONE || cfg->body > LBM_BODY_MASK)
        return set_err(nullptr, LBM_ERR_INVALID, "unknown body");
    if (ref_mode(cfg->bc_mode) && cfg->body == LBM_BODY_MASK)
        return set_err(nullptr, LBM_ERR_INVALID, "reference mode takes its body from the sphere/cylinder formula");
    if (cfg->inlet < LBM_INLET_UNIFORM || cfg->inlet > LBM_INLET_PROFILE)
        return set_err(nullptr, LBM_ERR_INVALID, "unknown inlet");
    if (cfg->ncoef < 0 || cfg->ncoef > 8) return set_err(nullptr, LBM_ERR_INVALID, "ncoef must be in [0, 8]");
    if (cfg->force_interval < 0) return set_err(nullptr, LBM_ERR_INVALID, "force_interval must be >= 0");
    const int nxp = (cfg->nx + 63) / 64 * 64;
    const size_t plane = (size_t)nxp * cfg->ny * nz;
    if (plane >= (size_t)1 << 32) return set_err(nullptr, LBM_ERR_INVALID, "domain too large: the padded node count nxp*ny*nz must stay below 2^32");

    lbm_ctx *c = new lbm_ctx();
    c->cfg = *cfg;
    c->cfg.nz = nz;
    c->Q = Q;
    c->D = Q == 9 ? 2 : 3;
    c->prec = cfg->precision;
    c->geo.nx = cfg->nx;
    c->geo.ny = cfg->ny;
    c->geo.nz = nz;
    c->geo.nxp = nxp;
    c->geo.plane = plane;
    // populations row-interleaved (the Q direction rows of a row side by side); the layout with
    // one plane per direction (LBM_DIRECTION_PLANES) is kept for comparison, same results
    const bool planes = std::getenv("LBM_DIRECTION_PLANES") != nullptr;
    c->geo.ds = planes ? plane : (size_t)nxp;
    c->geo.rs = planes ? (size_t)nxp : (size_t)Q * nxp;
    const bool plain = cfg->bc_mode == LBM_BC_PLAIN;
    c->geo.xper = plain && cfg->periodic[0];
    c->geo.yper = plain && cfg->periodic[1];
    c->geo.zper = plain && Q == 27 && cfg->periodic[2];
    c->N = (size_t)cfg->nx * cfg->ny * nz;
    // nodes per lane of the fused kernel: one 4-byte access per direction for two FP16/BF16
    // nodes; FP32 and FP64 keep one node per lane (measured fastest on the 512x256x256 case).
    // LBM_NODES_PER_LANE overrides it among the instantiated widths (FP32: 1, 2; FP16/BF16: 2, 4).
    c->V = (c->prec == LBM_FP16S || c->prec == LBM_BF16S) ? 2 : 1;
    if (const char *env = std::getenv("LBM_NODES_PER_LANE")) {
        const int v = std::atoi(env);
        const bool ok = c->prec == LBM_FP32 ? (v == 1 || v == 2) : c->prec != LBM_FP64 && (v == 2 || v == 4);
        if (ok) c->V = v;
    }
    // 16-bit storage with two nodes per lane pulls with one (2-byte aligned) dword load per
    // direction; LBM_ALIGNED_PULL selects aligned loads with lane shifts instead (same
    // results; faster with one plane per direction, slower with the row-interleaved layout).
    c->aligned_pull = std::getenv("LBM_ALIGNED_PULL") != nullptr;
    // Blocks of the packed kernel resident per CU (256-lane blocks, row-interleaved populations).
    // D3Q27, measured on the 512x256x256 cylinder: FP32 fastest at 3 (3 waves per SIMD: 71.5 %
    // of HBM peak per step against 70.8 % uncapped), FP16/BF16 at 5 (steps 6-25 of the seeded
    // start: 0.647 ms per step against 0.664 at 4 and 0.649 at 6; later steps equal at 4 and 5).
    // D2Q9, measured on an 8192x4096 disk: FP32 fastest uncapped (78.5 % against 69.3 % at 3),
    // FP16/BF16 at 6 (73.5 % against 71.7 % uncapped and 67.4 % at 4).  The cap is a dynamic LDS
    // allocation the kernel does not use.  LBM_FAST_BLOCKS_PER_CU overrides it (0: no cap).
    // Rows of fewer than 256 lanes (narrow domains) keep the same waves per CU.
    // Reference mode's fused kernel (FP64, D3Q27): fastest at 3 (512x128x128 sphere: 0.730 ms per
    // step against 0.767 uncapped, 0.761 at 2, 0.751 at 4); the plain FP64 kernel gains nothing
    // (0.628-0.630 ms at none, 3, 4; 0.701 at 2), so it stays uncapped.
    int bpc = 0;
    if (c->prec == LBM_FP32) bpc = Q == 27 ? 3 : 0;
    else if (c->prec != LBM_FP64) bpc = Q == 27 ? 5 : 6;
    else if (ref_mode(cfg->bc_mode) && Q == 27) bpc = 3;
    if (const char *env = std::getenv("LBM_FAST_BLOCKS_PER_CU")) bpc = std::max(0, std::atoi(env));
    hipError_t e = hipSuccess;
    if (bpc > 0) {
        int dev = 0, lds = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
        if (e != hipSuccess) {
            set_err(nullptr, LBM_ERR_HIP, std::string("hipDeviceGetAttribute: ") + hipGetErrorString(e));
            delete c;
            return LBM_ERR_HIP;
        }
        // the count is of 4-wave (256-lane) blocks: rows narrower than 256 lanes launch smaller
        // blocks, and proportionally more of them may stay resident (the plain kernels); the
        // reference mode's fused kernel always launches 256-lane blocks
        const int lanes = ref_mode(cfg->bc_mode) ? 256 : nxp / c->V, wpb = (std::min(lanes, 256) + 63) / 64;
        const int blocks = std::max(1, bpc * 4 / wpb);
        // midway between lds/(blocks+1) and lds/blocks, so allocation granularity cannot tip the count
        const unsigned per = (unsigned)(2LL * lds / (2 * blocks + 1)) / 512u * 512u;
        if (per > (unsigned)lds / (unsigned)(blocks + 1) && per <= 65536u) c->fast_lds = per;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_err(nullptr, LBM_ERR_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
        delete c;
        return LBM_ERR_HIP;
    }
    // lbm_get_fields_noalloc's staging array on this device (its own error, reported by that call)
    if (noalloc_setup(&c->noalloc_scratch, c->noalloc_err) != LBM_OK) c->noalloc_scratch = nullptr;
    *out = c;
    return LBM_OK;
}

int snapshot_wait(lbm_ctx *c);

extern "C" void lbm_destroy(lbm_ctx *c) {
    if (!c) return;
    snapshot_wait(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int lbm_set_mask(lbm_ctx *c, const uint8_t *solid) {
    if (!c || !solid) return set_err(c, LBM_ERR_INVALID, "null argument");
    if (c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_set_mask after lbm_initialize");
    if (c->cfg.body != LBM_BODY_MASK) return set_err(c, LBM_ERR_INVALID, "lbm_set_mask requires body = LBM_BODY_MASK");
    c->solid.assign(solid, solid + c->N);
    return LBM_OK;
}

extern "C" int lbm_set_solid_fraction(lbm_ctx *c, const double *s) {
    if (!c || !s) return set_err(c, LBM_ERR_INVALID, "null argument");
    if (c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_set_solid_fraction after lbm_initialize");
    if (c->cfg.bc_mode != LBM_BC_PLAIN) return set_err(c, LBM_ERR_INVALID, "partial bounce-back is a plain-mode feature");
    for (size_t n = 0; n < c->N; n++)
        if (!(s[n] >= 0.0 && s[n] <= 1.0)) return set_err(c, LBM_ERR_INVALID, "solid fraction outside [0, 1]");
    c->sfrac.assign(s, s + c->N);
    return LBM_OK;
}

extern "C" int lbm_set_inlet_profile(lbm_ctx *c, const double *uin) {
    if (!c || !uin) return set_err(c, LBM_ERR_INVALID, "null argument");
    if (c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_set_inlet_profile after lbm_initialize");
    c->uin.assign(uin, uin + (size_t)c->geo.ny * c->geo.nz);
    c->uin_set = true;
    return LBM_OK;
}

extern "C" int lbm_set_perturbation(lbm_ctx *c, const double *xi) {
    if (!c || !xi) return set_err(c, LBM_ERR_INVALID, "null argument");
    if (c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_set_perturbation after lbm_initialize");
    c->xi.assign(xi, xi + c->N);
    return LBM_OK;
}

static void ref_extents(const lbm_ctx *c, RefGeo &r) {
    r.Q = c->Q;
    r.D = c->D;
    r.g = c->D == 3 ? 3 : 2;   // ghost width: 3D MPI+KOKKOS+MPIIO/lbm.hpp:22, MPI Kokkos/lbm.hpp:79-80
    r.glx = c->geo.nx;
    r.gly = c->geo.ny;
    r.glz = c->geo.nz;
    r.lx = r.glx + 2 * r.g;
    r.ly = r.gly + 2 * r.g;
    r.lz = c->D == 3 ? r.glz + 2 * r.g : 1;
    // 3D rows padded to 32 doubles: a row of lx = nx + 6 doubles then starts on a 256-byte
    // boundary (about 15 % faster at nx = 512).  2D keeps px = lx, since its wall fills read
    // one column past a row end into the next row exactly as the reference's flat index does.
    r.px = c->D == 3 ? (r.lx + 31) / 32 * 32 : r.lx;
    r.plane = (size_t)r.px * r.ly * r.lz;
    // populations: 3D row-interleaved as in the plain path (every 3D access stays inside its
    // row: i +- 2 lies in [0, lx) at every node a fill, usrbc or pull touches), 2D one plane
    // per direction (the flat-index wrap above needs consecutive rows of one direction);
    // LBM_DIRECTION_PLANES selects planes in 3D too
    const bool planes = c->D == 2 || std::getenv("LBM_DIRECTION_PLANES") != nullptr;
    r.ds = planes ? r.plane : (size_t)r.px;
    r.rs = planes ? (size_t)r.px : (size_t)r.Q * r.px;
    r.tau0 = c->cfg.tau0;
    r.u0 = c->cfg.u0;
    r.variant = c->cfg.bc_mode == LBM_BC_REFERENCE_SN ? 1 : 0;
    r.lds = c->fast_lds;
}

extern "C" int lbm_set_ran(lbm_ctx *c, const int32_t *ran, size_t count) {
    if (!c || !ran) return set_err(c, LBM_ERR_INVALID, "null argument");
    if (c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_set_ran after lbm_initialize");
    if (!ref_mode(c->cfg.bc_mode)) return set_err(c, LBM_ERR_INVALID, "ran is the reference-mode perturbation");
    RefGeo r;
    ref_extents(c, r);
    if (count != (size_t)r.lx * r.ly * r.lz) return set_err(c, LBM_ERR_INVALID, "ran must cover the ghost-padded box lx*ly*lz");
    for (size_t n = 0; n < count; n++)
        if (ran[n] < 0 || ran[n] > 99) return set_err(c, LBM_ERR_INVALID, "ran values must lie in [0, 100)");
    c->ran.assign(ran, ran + count);
    return LBM_OK;
}

// ---- host-side geometry ----
static double legendre(int k, double x) {
    double p0 = 1.0, p1 = x;
    if (k == 0) return p0;
    for (int n = 1; n < k; n++) {
        const double p2 = ((2.0 * n + 1.0) * x * p1 - (double)n * p0) / (double)(n + 1);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// inlet velocity u_x(y, z); walls sit half a spacing outside the first/last node
static void inlet_profile(const lbm_ctx *c, std::vector<double> &uin) {
    static const int LY[8] = {0, 1, 0, 2, 1, 0, 3, 0}, LZ[8] = {0, 0, 1, 0, 1, 2, 0, 3};
    const int ny = c->geo.ny, nz = c->geo.nz, D = c->D;
    uin.assign((size_t)ny * nz, 0.0);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++) {
            const double ey = ((double)y + 0.5) / (double)ny, ez = ((double)z + 0.5) / (double)nz;
            const double by = 4.0 * ey * (1.0 - ey), bz = D == 3 ? 4.0 * ez * (1.0 - ez) : 1.0;
            double u;
            if (c->cfg.inlet == LBM_INLET_PARABOLIC) {
                u = c->cfg.u0 * by * bz;
            } else if (c->cfg.inlet == LBM_INLET_LEGENDRE) {
                double sum = 0.0;
                for (int k = 0; k < c->cfg.ncoef; k++) {
                    const double phi = D == 3 ? legendre(LY[k], 2.0 * ey - 1.0) * legendre(LZ[k], 2.0 * ez - 1.0)
                                              : legendre(k, 2.0 * ey - 1.0);
                    sum += c->cfg.coef[k] * phi;
                }
                u = by * bz * sum;
            } else {
                u = c->cfg.u0;
            }
            uin[y + (size_t)ny * z] = u;
        }
}

static bool body_inside(const lbm_config &cfg, int D, int x, int y, int z) {
    if (cfg.body != LBM_BODY_SPHERE && cfg.body != LBM_BODY_ZCYL) return false;
    const double dx = (double)x - cfg.cx, dy = (double)y - cfg.cy;
    double r = dx * dx + dy * dy;
    if (D == 3 && cfg.body == LBM_BODY_SPHERE) {
        const double dz = (double)z - cfg.cz;
        r = r + dz * dz;
    }
    return r <= cfg.r2;
}

static int grow_hist(lbm_ctx *c, int need) {
    if (need <= c->hist_cap) return LBM_OK;
    int cap = std::max(need, std::max(1024, 2 * c->hist_cap));
    double *h = nullptr;
    HIP_TRY(c, hipMalloc(&h, (size_t)cap * 4 * sizeof(double)));
    if (c->hist) {
        HIP_TRY(c, hipMemcpyAsync(h, c->hist, (size_t)c->nrec * 4 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->hist);
    }
    c->hist = h;
    c->hist_cap = cap;
    return LBM_OK;
}

FusedArgs ctx_fused_args(const lbm_ctx *c) {
    FusedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.g = c->geo;
    a.tau = c->cfg.tau0 + 0.5;
    a.rtau = 1.0 / a.tau;
    a.omega = (float)(1.0 / (c->cfg.tau0 + 0.5));
    // 16-bit formats: offset o_i = feq_i(0, (u_ref, 0, 0)) and o_i / w_i
    std::vector<int> e(3 * c->Q), opp(c->Q);
    std::vector<double> w(c->Q);
    lbm_lattice_table(c->Q, e.data(), w.data(), opp.data());
    const double U = c->cfg.store_uref;
    for (int i = 0; i < c->Q; i++) {
        const double eu = e[3 * i] * U;
        const double o = w[i] * (3.0 * eu + 4.5 * eu * eu - 1.5 * U * U);
        a.sp.off[i] = (float)o;
        a.sp.offw[i] = (float)(o / w[i]);
    }
    return a;
}

static int init_plain(lbm_ctx *c) {
    const Geo &g = c->geo;
    const int Q = c->Q, D = c->D;
    // solid mask: the body formula (the reference's "usr", 3D MPI+KOKKOS+MPIIO/lbm.cpp:238) or a user mask
    if (c->cfg.body == LBM_BODY_MASK) {
        if (c->solid.size() != c->N) return set_err(c, LBM_ERR_STATE, "body = LBM_BODY_MASK needs lbm_set_mask first");
    } else {
        c->solid.assign(c->N, 0);
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++)
                    c->solid[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] = body_inside(c->cfg, D, x, y, z);
    }
    if (!c->uin_set) inlet_profile(c, c->uin);
    // node flags (padded layout); padding columns are marked solid
    std::vector<int> e(3 * Q), opp(Q);
    lbm_lattice_table(Q, e.data(), nullptr, opp.data());
    std::vector<uint8_t> flags(g.plane, FL_SOLID);
    auto solid_at = [&](int x, int y, int z) -> bool {
        if (g.xper) x = (x + g.nx) % g.nx;
        if (g.yper) y = (y + g.ny) % g.ny;
        if (g.zper) z = (z + g.nz) % g.nz;
        if (x < 0 || x >= g.nx || y < 0 || y >= g.ny || z < 0 || z >= g.nz) return false;
        return c->solid[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] != 0;
    };
    c->solid_nodes = c->near_nodes = 0;
    for (int z = 0; z < g.nz; z++)
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
                if (solid_at(x, y, z)) {
                    flags[node] = FL_SOLID;
                    c->solid_nodes++;
                    continue;
                }
                bool near = false;
                for (int i = 1; i < Q && !near; i++) near = solid_at(x - e[3 * i], y - e[3 * i + 1], z - e[3 * i + 2]);
                flags[node] = near ? FL_NEAR : FL_FLUID;
                c->near_nodes += near;
            }
    const int V = c->V;
    const bool sf = !c->sfrac.empty();
    auto forced = [&](int x, int y, int z) {   // a node whose links or solid fraction carry force
        if (x >= g.nx) return false;
        const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
        return flags[node] == FL_NEAR ||
               (sf && flags[node] != FL_SOLID && c->sfrac[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] > 0.0);
    };
    // per-node kernels over every node (FP64 k_fused, dense mode k_dense_v; one node per lane):
    // a force partial slot per block holding a forced node
    std::vector<int> slot;
    c->nslots = 0;
    auto block_slots = [&]() {
        const Shape s1 = c->prec == LBM_FP64 ? fp64_shape(g, Q) : lane_shape(g, 1);
        slot.assign((size_t)s1.gx * g.ny * g.nz, -1);
        c->nslots = 0;
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int bxi = 0; bxi < s1.gx; bxi++) {
                    bool any = false;
                    for (int x = bxi * s1.bx; x < std::min(g.nx, (bxi + 1) * s1.bx) && !any; x++) any = forced(x, y, z);
                    if (any) slot[(size_t)bxi + (size_t)s1.gx * ((size_t)y + (size_t)g.ny * z)] = c->nslots++;
                }
    };
    if (c->prec == LBM_FP64) block_slots();
    // FP32/16-bit: per fluid node (row-major padded indices), either
    //  * the link list: mask neighbours whose half-way bounce-back the packed kernel applies
    //    in its pull, from per-row windows of link words and force slots (below), or
    //  * the fix list, recomputed per node after the packed path: nodes with a solid fraction,
    //    the periodic x edges, the last node when nx is not a multiple of the lane width V
    //    (its lane reaches into the padding), and mask neighbours in the inlet or outlet lane
    //    (whose packs the packed kernel shifts by one node) or, when opposite directions have
    //    different storage offsets (16-bit storage with store_uref != 0), every mask neighbour
    //    (the packed kernel substitutes stored values, not decoded ones).
    // LBM_FORCE_SLOW_PATH=1 puts every fluid node on the fix list.  Force partial slots go to
    // the forced nodes in node order across both lists, so the force sum does not depend on
    // which list a node is on.
    if (c->prec != LBM_FP64) {
        const char *force_slow = std::getenv("LBM_FORCE_SLOW_PATH");
        const bool all_slow = force_slow && force_slow[0] == '1';
        const bool same_offsets = c->prec == LBM_FP32 || c->cfg.store_uref == 0.0;
        std::vector<uint32_t> fix, fixlk, link, linklk;
        std::vector<int> lslot;
        size_t fluid = 0;
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
                    if (flags[node] == FL_SOLID) continue;
                    fluid++;
                    const bool f = forced(x, y, z), near = flags[node] == FL_NEAR;
                    const bool sfr = sf && c->sfrac[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] > 0.0;
                    const bool edge = (g.xper && (x == 0 || x == g.nx - 1)) || (x == g.nx - 1 && g.nx % V != 0);
                    const int x0 = x / V * V;   // first node of the packed kernel's lane
                    const bool end_lane = x0 == 0 || x0 + V >= g.nx;
                    const bool to_fix = all_slow || sfr || edge || (near && (end_lane || !same_offsets));
                    if (!to_fix && !near) continue;
                    // link word: the per-node pull's classification (Zou-He unknowns and wall
                    // links first, outlet clamp, periodic wrap), bit i when the source is solid
                    uint32_t bits = 0;
                    if (near)
                        for (int i = 1; i < Q; i++) {
                            const int ex = e[3 * i], ey = e[3 * i + 1], ez = e[3 * i + 2];
                            if (ex > 0 && x == 0 && !g.xper) continue;
                            int sx = x - ex;
                            const int sy = y - ey, sz = z - ez;
                            if ((sy < 0 || sy >= g.ny) && !g.yper) continue;
                            if ((sz < 0 || sz >= g.nz) && !g.zper) continue;
                            if (sx == g.nx && !g.xper) sx = g.nx - 1;
                            if (solid_at(sx, sy, sz)) bits |= 1u << i;
                        }
                    const int k = f ? c->nslots++ : -1;
                    if (to_fix) {
                        fix.push_back((uint32_t)node);
                        fixlk.push_back(bits);
                        slot.push_back(k);
                    } else {
                        link.push_back((uint32_t)node);
                        linklk.push_back(bits);
                        lslot.push_back(k);
                    }
                }
        // Dense mode: one per-node pass over every node instead of the copies, the packed pass
        // and the fix-up of the list, chosen
        //  * when the fix list holds more than a quarter of the fluid nodes (a solid fraction
        //    over a noisy field: half of them on the 192x96x96 adjoint workload, where the
        //    list path took 67 + 100 us per step and the dense pass 67 us), or
        //  * when a population buffer is at most 16 MB: such steps are launch-bound, and one
        //    launch fewer per step counts (D2Q9 512x128 with the drag every step: 6.7 / 6.4 us
        //    per step in FP32 / FP16 against 8.7 / 8.8; on D3Q27 512x128x128 the dense pass
        //    is as fast in FP32 but 1.5 times slower in FP16, so large domains keep the list).
        // LBM_DENSE=1 / 0 forces the choice.
        const size_t pop_bytes = (size_t)(c->prec == LBM_FP32 ? 4 : 2) * Q * g.plane;
        const bool small = pop_bytes <= ((size_t)16 << 20);
        const char *dense_env = std::getenv("LBM_DENSE");
        c->dense = dense_env ? dense_env[0] == '1' : (!all_slow && (fix.size() * 4 > fluid || small));
        if (c->dense) {
            fix.clear();
            fixlk.clear();
            link.clear();
            linklk.clear();
            lslot.clear();
            block_slots();
        }
        if (fix.size() > (size_t)INT32_MAX || link.size() > (size_t)INT32_MAX)
            return set_err(c, LBM_ERR_INVALID, "too many boundary nodes");
        c->nfix = (int)fix.size();
        c->nlink = (int)link.size();
        if (c->nfix) {
            HIP_TRY(c, hipMalloc(&c->fix, sizeof(uint32_t) * fix.size()));
            HIP_TRY(c, hipMemcpy(c->fix, fix.data(), sizeof(uint32_t) * fix.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMalloc(&c->fixlk, sizeof(uint32_t) * fixlk.size()));
            HIP_TRY(c, hipMemcpy(c->fixlk, fixlk.data(), sizeof(uint32_t) * fixlk.size(), hipMemcpyHostToDevice));
        }
        // Per-row windows of the link list for the packed kernel: row r's lanes [t0, t0 + nl)
        // cover its linked nodes; the window's nl*V nodes hold their link words (0 for nodes
        // without links) from element `base` of wlk.  The list itself serves force steps.
        if (c->nlink) {
            std::vector<int4> win((size_t)g.ny * g.nz, make_int4(0, 0, 0, 0));
            std::vector<uint32_t> wlk;
            for (size_t j = 0; j < link.size();) {
                const uint32_t row = link[j] / (uint32_t)g.nxp;
                size_t k = j;
                while (k < link.size() && link[k] / (uint32_t)g.nxp == row) k++;   // the row's entries (node order)
                const int t0 = (int)(link[j] - row * (uint32_t)g.nxp) / V;
                const int t1 = (int)(link[k - 1] - row * (uint32_t)g.nxp) / V;
                const size_t base = wlk.size();
                if (base + (size_t)(t1 - t0 + 1) * V > (size_t)INT32_MAX)
                    return set_err(c, LBM_ERR_INVALID, "too many mask links");
                wlk.resize(base + (size_t)(t1 - t0 + 1) * V, 0u);
                for (size_t m = j; m < k; m++)
                    wlk[base + (link[m] - row * (uint32_t)g.nxp) - (size_t)t0 * V] = linklk[m];
                win[row] = make_int4(t0, t1 - t0 + 1, (int)base, 0);
                j = k;
            }
            HIP_TRY(c, hipMalloc(&c->rowwin, sizeof(int4) * win.size()));
            HIP_TRY(c, hipMemcpy(c->rowwin, win.data(), sizeof(int4) * win.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMalloc(&c->wlk, sizeof(uint32_t) * wlk.size()));
            HIP_TRY(c, hipMemcpy(c->wlk, wlk.data(), sizeof(uint32_t) * wlk.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMalloc(&c->link, sizeof(uint32_t) * link.size()));
            HIP_TRY(c, hipMemcpy(c->link, link.data(), sizeof(uint32_t) * link.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMalloc(&c->linklk, sizeof(uint32_t) * linklk.size()));
            HIP_TRY(c, hipMemcpy(c->linklk, linklk.data(), sizeof(uint32_t) * linklk.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMalloc(&c->link_slot, sizeof(int) * lslot.size()));
            HIP_TRY(c, hipMemcpy(c->link_slot, lslot.data(), sizeof(int) * lslot.size(), hipMemcpyHostToDevice));
        }
    }
    // D2Q9 FP64 on a small domain, where a step costs about a dependent launch: several steps
    // per launch (temporal blocking, small_kernels.hip), for walls, inlet, outlet and a mask,
    // when the population buffer is at most 24 MB (measured against one launch per step: 2.3x
    // at 3 MB, 1.08x at 17.6 MB, even at 34 MB, slower above); LBM_TEMPORAL_BLOCKING=0 / 1
    // forbids / allows it regardless of the size.
    {
        const char *tbe = std::getenv("LBM_TEMPORAL_BLOCKING");
        const bool eligible = Q == 9 && c->prec == LBM_FP64 && !g.xper && !g.yper && !sf;
        const bool small = (size_t)8 * Q * g.plane <= ((size_t)24 << 20);
        int dev = 0, lds_max = 0, cus = 0;
        if (eligible && (tbe ? tbe[0] == '1' : small) && hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            tb_prepare() == hipSuccess) {
            const int K = tb_depth(), T = tb_tile(g, K, cus);
            if ((size_t)lds_max >= tb_lds_bytes(T) + 4096 + 512) {   // + the kernel's static force records
                c->tb_k = K;
                c->tb_t = T;
                int tiles_x = 0;
                const int nt = tb_tiles(g, T, K, &tiles_x);
                HIP_TRY(c, hipMalloc(&c->tb_partials, sizeof(double) * 3 * (size_t)nt * tb_records(K)));
                const std::vector<uint16_t> lk = tb_links(g, flags);
                HIP_TRY(c, hipMalloc(&c->tb_lk, sizeof(uint16_t) * lk.size()));
                HIP_TRY(c, hipMemcpy(c->tb_lk, lk.data(), sizeof(uint16_t) * lk.size(), hipMemcpyHostToDevice));
            }
        }
    }
    const size_t esz = c->prec == LBM_FP64 ? 8 : c->prec == LBM_FP32 ? 4 : 2;
    HIP_TRY(c, hipMalloc(&c->A, esz * Q * g.plane));
    HIP_TRY(c, hipMalloc(&c->B, esz * Q * g.plane));
    HIP_TRY(c, hipMalloc(&c->flags, g.plane));
    HIP_TRY(c, hipMalloc(&c->duin, sizeof(double) * c->uin.size()));
    HIP_TRY(c, hipMalloc(&c->macros, sizeof(double) * 4 * g.plane));
    HIP_TRY(c, hipMalloc(&c->block_slot, sizeof(int) * std::max<size_t>(slot.size(), 1)));
    HIP_TRY(c, hipMalloc(&c->partials, sizeof(double) * 3 * (std::max(c->nslots, 1) + force_chunks(c->nslots))));
    HIP_TRY(c, hipMemcpyAsync(c->flags, flags.data(), g.plane, hipMemcpyHostToDevice, c->stream));
    {   // rows whose nodes are all fluid with no solid neighbour (the adjoint's straight-line rows;
        // LBM_ADJ_GENERAL_PATH=1 sends every row through the general path, for tests)
        std::vector<uint8_t> clean((size_t)g.ny * g.nz, std::getenv("LBM_ADJ_GENERAL_PATH") ? 0 : 1);
        for (size_t r = 0; r < clean.size(); r++)
            for (int x = 0; x < g.nx && clean[r]; x++) clean[r] = flags[(size_t)x + (size_t)g.nxp * r] == FL_FLUID;
        HIP_TRY(c, hipMalloc(&c->rowclean, clean.size()));
        HIP_TRY(c, hipMemcpy(c->rowclean, clean.data(), clean.size(), hipMemcpyHostToDevice));
        c->adj_all_fast = true;
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++) {
                const bool pedge = (g.yper && (y == 0 || y == g.ny - 1)) || (Q == 27 && g.zper && (z == 0 || z == g.nz - 1));
                if (!clean[(size_t)y + (size_t)g.ny * z] || pedge) c->adj_all_fast = false;
            }
    }
    HIP_TRY(c, hipMemcpyAsync(c->duin, c->uin.data(), sizeof(double) * c->uin.size(), hipMemcpyHostToDevice, c->stream));
    if (!slot.empty())
        HIP_TRY(c, hipMemcpy(c->block_slot, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice));
    if (sf) {
        const size_t ssz = c->prec == LBM_FP64 ? 8 : 4;
        std::vector<char> buf(ssz * g.plane, 0);
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
                    const double v = c->sfrac[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)];
                    if (ssz == 8) ((double *)buf.data())[node] = v;
                    else ((float *)buf.data())[node] = (float)v;
                }
        HIP_TRY(c, hipMalloc(&c->dsfrac, buf.size()));
        HIP_TRY(c, hipMemcpyAsync(c->dsfrac, buf.data(), buf.size(), hipMemcpyHostToDevice, c->stream));
    }
    double *dxi = nullptr;
    if (!c->xi.empty()) {
        HIP_TRY(c, hipMalloc(&dxi, sizeof(double) * c->N));
        HIP_TRY(c, hipMemcpyAsync(dxi, c->xi.data(), sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
    }
    const FusedArgs a = ctx_fused_args(c);
    HIP_TRY(c, plain_init(Q, c->prec, c->A, c->B, c->macros, c->flags, c->duin, dxi, c->cfg.perturb, c->cfg.seed, a, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (dxi) (void)hipFree(dxi);
    return LBM_OK;
}

static int init_ref(lbm_ctx *c) {
    RefGeo &r = c->rg;
    ref_extents(c, r);
    const int Q = c->Q;
    HIP_TRY(c, hipMalloc(&c->rf, sizeof(double) * Q * r.plane));
    HIP_TRY(c, hipMalloc(&c->rft, sizeof(double) * Q * r.plane));
    HIP_TRY(c, hipMalloc(&c->rmac, sizeof(double) * 4 * r.plane));
    HIP_TRY(c, hipMalloc(&c->rusr, r.plane));
    HIP_TRY(c, hipMalloc(&c->rfpre, sizeof(double) * Q * r.plane));
    const int nprof = c->D == 3 ? r.glz : r.gly;
    std::vector<double> prof(nprof);
    for (int k = 0; k < nprof; k++)   // u0*4*c*(n-1-c)/(n-1)^2, lbm.cpp:405 / MPI Kokkos/lbm.cpp:446
        prof[k] = c->cfg.u0 * 4.0 * (double)k * (double)(nprof - 1 - k) / ((double)(nprof - 1) * (double)(nprof - 1));
    HIP_TRY(c, hipMalloc(&c->rprof, sizeof(double) * nprof));
    HIP_TRY(c, hipMemcpyAsync(c->rprof, prof.data(), sizeof(double) * nprof, hipMemcpyHostToDevice, c->stream));
    int32_t *dran = nullptr;
    if (!c->ran.empty()) {
        HIP_TRY(c, hipMalloc(&dran, sizeof(int32_t) * c->ran.size()));   // unpadded lx*ly*lz
        HIP_TRY(c, hipMemcpyAsync(dran, c->ran.data(), sizeof(int32_t) * c->ran.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(c, ref_init(r, c->rf, c->rft, c->rmac, c->rusr, dran, c->cfg.body, c->cfg.cx, c->cfg.cy, c->cfg.cz,
                        c->cfg.r2, c->cfg.seed, c->stream));
    // before the first step the reference's f is the initial equilibrium
    HIP_TRY(c, hipMemcpyAsync(c->rfpre, c->rf, sizeof(double) * Q * r.plane, hipMemcpyDeviceToDevice, c->stream));
    std::vector<uint8_t> usr(r.plane);
    HIP_TRY(c, hipMemcpyAsync(usr.data(), c->rusr, r.plane, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (dran) (void)hipFree(dran);
    // usrbc works in place; reject masks for which its result would depend on thread order
    // (a solid node, a fluid gap of one node and another solid node along a link)
    std::vector<int> e(3 * Q);
    lbm_lattice_table(Q, e.data(), nullptr, nullptr);
    const int g = r.g, klo = c->D == 3 ? g : 0, khi = c->D == 3 ? r.lz - g : 1;
    c->solid_nodes = 0;
    std::vector<uint32_t> surf;   // usrbc's work list: interior solid nodes with a fluid neighbour
    for (int k = klo; k < khi; k++)
        for (int j = g; j < r.ly - g; j++)
            for (int i = g; i < r.lx - g; i++) {
                const long n = i + (long)r.px * (j + (long)r.ly * k);
                if (usr[n] != 0) continue;
                c->solid_nodes++;
                bool wet = false;
                for (int q = 1; q < Q; q++) {
                    const long d = e[3 * q] + (long)r.px * (e[3 * q + 1] + (long)r.ly * e[3 * q + 2]);
                    if (usr[n + d] == 1 && usr[n + 2 * d] == 0)
                        return set_err(c, LBM_ERR_INVALID, "reference mode: body has a one-node fluid gap along a lattice link (usrbc would race)");
                    wet = wet || usr[n + d] == 1;
                }
                if (wet) surf.push_back((uint32_t)n);
            }
    // 3D only: the 2D box (65 536 nodes by default) is launch-bound, and its step measured 7 %
    // slower with the list than with the interior sweep
    if (c->D == 3 && r.plane <= UINT32_MAX) {
        c->nsurf = (int)surf.size();
        HIP_TRY(c, hipMalloc(&c->rsurf, sizeof(uint32_t) * std::max<size_t>(surf.size(), 1)));
        if (!surf.empty())
            HIP_TRY(c, hipMemcpy(c->rsurf, surf.data(), sizeof(uint32_t) * surf.size(), hipMemcpyHostToDevice));
    }
    const int nblk = ref_mem_blocks(r);
    HIP_TRY(c, hipMalloc(&c->partials, sizeof(double) * 3 * nblk));
    c->nslots = nblk;
    return LBM_OK;
}

extern "C" int lbm_initialize(lbm_ctx *c) {
    if (!c) return LBM_ERR_INVALID;
    if (c->initialized) return set_err(c, LBM_ERR_STATE, "already initialised");
    int rc = c->cfg.bc_mode == LBM_BC_PLAIN ? init_plain(c) : init_ref(c);
    if (rc == LBM_OK) {
        const hipError_t e = hipMalloc(&c->dbad, sizeof(unsigned));
        if (e != hipSuccess) rc = set_err(c, LBM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    if (rc != LBM_OK) {
        free_all(c);
        return rc;
    }
    c->initialized = true;
    c->step = 0;
    c->stage = 0;
    c->nrec = c->nrec_checked = 0;
    c->fields_fresh = true;   // the initial macros
    c->replay_ok = false;
    return LBM_OK;
}

StepBufs step_bufs(const lbm_ctx *c, const void *A, void *B) {
    StepBufs b;
    b.A = A;
    b.B = B;
    b.flags = c->flags;
    b.fix = c->fix;
    b.fixlk = c->fixlk;
    b.nfix = c->nfix;
    b.rowwin = c->rowwin;
    b.wlk = c->wlk;
    b.link = c->link;
    b.linklk = c->linklk;
    b.link_slot = c->link_slot;
    b.nlink = c->nlink;
    b.sfrac = c->dsfrac;
    b.uin = c->duin;
    b.block_slot = c->block_slot;
    b.dense = c->dense;
    b.partials = c->partials;
    b.macros = c->macros;
    b.aligned_pull = c->aligned_pull;
    b.fast_lds = c->fast_lds;
    b.ev_bulk[0] = b.ev_bulk[1] = nullptr;
    return b;
}

// Reference mode, after the fills: usrbc and, on force steps, the momentum exchange.  In 2D
// (launch-bound, the drag every step by default) one pass does both; in 3D usrbc runs over its
// list and the exchange separately.
static hipError_t ref_boundary_pass(lbm_ctx *c, bool force) {
    if (force && c->D == 2)
        return ref_mem(c->rg, c->rf, c->rusr, c->partials, c->nslots, c->hist, c->nrec++, c->step, true, c->stream);
    hipError_t e = ref_usrbc(c->rg, c->rf, c->rusr, c->rsurf, c->nsurf, c->stream);
    if (e == hipSuccess && force)
        e = ref_mem(c->rg, c->rf, c->rusr, c->partials, c->nslots, c->hist, c->nrec++, c->step, false, c->stream);
    return e;
}

// evs: null, or 2n events recorded around each step's bulk kernel (plain mode)
static int step_n(lbm_ctx *c, int n, hipEvent_t *evs = nullptr) {
    if (c->stage != 0)
        return set_err(c, LBM_ERR_STATE, "a staged step is in progress: finish it with lbm_boundary, lbm_stream, lbm_update");
    const int fi = c->cfg.force_interval;
    const int nforce = fi > 0 ? (int)((c->step + n) / fi - c->step / fi) : 0;
    int rc = grow_hist(c, c->nrec + nforce + 1);
    if (rc != LBM_OK) return rc;
    if (ref_mode(c->cfg.bc_mode)) {
        for (int k = 0; k < n; k++) {
            c->step++;
            const bool force = fi > 0 && c->step % fi == 0;
            HIP_TRY(c, ref_fills(c->rg, c->rf, c->rft, c->stream));
            HIP_TRY(c, ref_boundary_pass(c, force));
            HIP_TRY(c, ref_fused(c->rg, c->rf, c->rft, nullptr, c->rusr, c->rprof, nullptr, c->stream));
            std::swap(c->rf, c->rft);
            c->fields_fresh = false;
            c->replay_ok = true;
        }
        return LBM_OK;
    }
    const FusedArgs a = ctx_fused_args(c);
    // temporally blocked launches for all but the last step (which stays a one-step launch, so
    // its input buffer can serve the fields on request); not under per-kernel timing
    // (their force records are summed together, when the partials' slots are full and before the
    // one-step launches, which write records of their own)
    int left = n;
    int pend = 0, rec0 = 0;
    long step1 = 0;
    while (c->tb_k > 0 && !evs && left > 1) {
        const int kk = std::min(c->tb_k, left - 1);
        if (pend + kk > tb_records(c->tb_k)) {
            HIP_TRY(c, tb_force_flush(c->tb_partials, a.g, c->tb_t, c->tb_k, c->hist, rec0, pend, step1, fi, c->stream));
            pend = 0;
        }
        if (pend == 0 && fi > 0) {
            rec0 = c->nrec;
            step1 = (c->step + fi) / fi * fi;   // the first force step after the current one
        }
        int nr = 0;
        HIP_TRY(c, tb_steps(c->A, c->B, c->flags, c->tb_lk, c->duin, c->tb_partials, pend, &nr, a, c->tb_t, c->tb_k, kk, c->step, fi,
                            c->stream));
        pend += nr;
        c->nrec += nr;
        c->step += kk;
        left -= kk;
        std::swap(c->A, c->B);
        c->fields_fresh = false;
        c->replay_ok = false;
    }
    HIP_TRY(c, tb_force_flush(c->tb_partials, a.g, c->tb_t, c->tb_k, c->hist, rec0, pend, step1, fi, c->stream));
    for (int k = n - left; k < n; k++) {
        c->step++;
        const bool force = fi > 0 && c->step % fi == 0;
        StepBufs b = step_bufs(c, c->A, c->B);
        if (evs) {
            b.ev_bulk[0] = evs[2 * k];
            b.ev_bulk[1] = evs[2 * k + 1];
        }
        HIP_TRY(c, plain_step(c->Q, c->prec, c->V, b, force, false, a, c->stream));
        if (force) HIP_TRY(c, force_reduce(c->partials, c->nslots, c->partials + 3 * (size_t)c->nslots, c->hist, c->nrec++, c->step, c->stream));
        std::swap(c->A, c->B);
        c->fields_fresh = false;
        c->replay_ok = true;
    }
    return LBM_OK;
}

// The moments of the last step (and, in reference mode, the populations after its Update) are
// computed on request: the steps write only populations, and the last step is run again from
// its input buffer, which no later work has touched, with the moments on.  The populations it
// writes are those already in the output buffer, bit for bit (same kernel, same input), so the
// state does not change; the cost is one step, paid only when fields are read.
int ensure_fields(lbm_ctx *c) {
    if (c->fields_fresh) return LBM_OK;
    if (!c->replay_ok) return set_err(c, LBM_ERR_STATE, "fields of the current state are not available");
    if (ref_mode(c->cfg.bc_mode)) {
        HIP_TRY(c, ref_fused(c->rg, c->rft, c->rf, c->rmac, c->rusr, c->rprof, c->rfpre, c->stream));
    } else {
        const FusedArgs a = ctx_fused_args(c);
        HIP_TRY(c, plain_step(c->Q, c->prec, c->V, step_bufs(c, c->B, c->A), false, true, a, c->stream));
    }
    c->fields_fresh = true;
    return LBM_OK;
}

extern "C" int lbm_step(lbm_ctx *c, int n) {
    if (!c) return LBM_ERR_INVALID;
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_step before lbm_initialize");
    if (n < 0) return set_err(c, LBM_ERR_INVALID, "negative step count");
    if (n == 0) return LBM_OK;
    return step_n(c, n);
}

// ---- staged step (reference mode) ----
// The fused step leaves rf holding the next step's collided populations (the reference's first
// Collision acts on f = feq and changes nothing, and every later one is applied right after
// Update), so lbm_collide only opens the step; lbm_boundary, lbm_stream and lbm_update then run
// the reference's Streaming() in two parts and its Update() (followed by the next Collision, in
// place), leaving the same state as lbm_step(ctx, 1).
static int staged(lbm_ctx *c, int want, const char *name) {
    if (!c) return LBM_ERR_INVALID;
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, std::string(name) + " before lbm_initialize");
    if (!ref_mode(c->cfg.bc_mode))
        return set_err(c, LBM_ERR_INVALID, std::string(name) + ": staged calls need a reference boundary mode "
                                                               "(the plain mode fuses the stages into one pass)");
    static const char *const order[4] = {"lbm_update (or nothing)", "lbm_collide", "lbm_boundary", "lbm_stream"};
    if (c->stage != want)
        return set_err(c, LBM_ERR_STATE, std::string(name) + " must follow " + order[want]);
    return LBM_OK;
}

extern "C" int lbm_collide(lbm_ctx *c) {
    int rc = staged(c, 0, "lbm_collide");
    if (rc != LBM_OK) return rc;
    // the staged fills write the ghost slots of both buffers, so the last fused step cannot be
    // rerun once the step is open: materialise its fields first
    rc = ensure_fields(c);
    if (rc != LBM_OK) return rc;
    c->replay_ok = false;
    rc = grow_hist(c, c->nrec + 2);
    if (rc != LBM_OK) return rc;
    c->step++;
    c->stage = 1;
    return LBM_OK;
}

extern "C" int lbm_boundary(lbm_ctx *c) {
    const int rc = staged(c, 1, "lbm_boundary");
    if (rc != LBM_OK) return rc;
    const int fi = c->cfg.force_interval;
    HIP_TRY(c, ref_fills(c->rg, c->rf, c->rft, c->stream));
    HIP_TRY(c, ref_boundary_pass(c, fi > 0 && c->step % fi == 0));
    c->stage = 2;
    return LBM_OK;
}

extern "C" int lbm_stream(lbm_ctx *c) {
    const int rc = staged(c, 2, "lbm_stream");
    if (rc != LBM_OK) return rc;
    HIP_TRY(c, ref_stream(c->rg, c->rf, c->rft, c->stream));
    std::swap(c->rf, c->rft);   // stream2: the ghost slots of both buffers are equal
    c->stage = 3;
    return LBM_OK;
}

extern "C" int lbm_update(lbm_ctx *c) {
    const int rc = staged(c, 3, "lbm_update");
    if (rc != LBM_OK) return rc;
    HIP_TRY(c, ref_update(c->rg, c->rf, c->rmac, c->rusr, c->rprof, c->rfpre, c->stream));
    c->stage = 0;
    c->fields_fresh = true;   // ref_update writes the moments and the populations after Update
    c->replay_ok = false;     // (it works in place: the fused kernel's input is gone)
    return LBM_OK;
}

extern "C" int lbm_step_timed(lbm_ctx *c, int n, double *ms) {
    if (!c || !ms) return LBM_ERR_INVALID;
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_step_timed before lbm_initialize");
    hipEvent_t e0, e1;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    HIP_TRY(c, hipEventRecord(e0, c->stream));
    int rc = n > 0 ? step_n(c, n) : LBM_OK;
    HIP_TRY(c, hipEventRecord(e1, c->stream));
    HIP_TRY(c, hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t;
    return rc;
}

extern "C" int lbm_step_kernel_timed(lbm_ctx *c, int n, double *ms_steps, double *ms_bulk) {
    if (!c || !ms_steps || !ms_bulk) return LBM_ERR_INVALID;
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_step_kernel_timed before lbm_initialize");
    if (c->cfg.bc_mode != LBM_BC_PLAIN) return set_err(c, LBM_ERR_INVALID, "lbm_step_kernel_timed times the plain mode");
    if (n <= 0 || n > 100000) return set_err(c, LBM_ERR_INVALID, "lbm_step_kernel_timed: n must be in [1, 100000]");
    std::vector<hipEvent_t> ev(2 * (size_t)n + 2, nullptr);
    auto release = [&]() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    for (hipEvent_t &e : ev) {
        const hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            e = nullptr;
            release();
            return set_err(c, LBM_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(r));
        }
    }
    hipError_t r = hipEventRecord(ev[2 * n], c->stream);
    int rc = r == hipSuccess ? step_n(c, n, ev.data()) : LBM_ERR_HIP;
    if (rc == LBM_OK) r = hipEventRecord(ev[2 * n + 1], c->stream);
    if (rc == LBM_OK && r == hipSuccess) r = hipEventSynchronize(ev[2 * n + 1]);
    float t = 0.f;
    double bulk = 0.0;
    if (rc == LBM_OK && r == hipSuccess) r = hipEventElapsedTime(&t, ev[2 * n], ev[2 * n + 1]);
    for (int k = 0; k < n && rc == LBM_OK && r == hipSuccess; k++) {
        float tk = 0.f;
        r = hipEventElapsedTime(&tk, ev[2 * k], ev[2 * k + 1]);
        bulk += tk;
    }
    release();
    if (rc != LBM_OK) return rc;
    if (r != hipSuccess) return set_err(c, LBM_ERR_HIP, std::string("lbm_step_kernel_timed: ") + hipGetErrorString(r));
    *ms_steps = t;
    *ms_bulk = bulk;
    return LBM_OK;
}

extern "C" int lbm_sync(lbm_ctx *c) {
    if (!c) return LBM_ERR_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->nrec > c->nrec_checked) {
        const int k = c->nrec - c->nrec_checked;
        std::vector<double> h(4 * (size_t)k);
        HIP_TRY(c, hipMemcpy(h.data(), c->hist + 4 * (size_t)c->nrec_checked, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        const int first = c->nrec_checked;
        c->nrec_checked = c->nrec;
        for (int r = 0; r < k; r++)
            for (int d = 1; d < 4; d++) {
                const double v = h[4 * (size_t)r + d];
                if (!std::isfinite(v) || (c->cfg.force_limit > 0.0 && std::fabs(v) > c->cfg.force_limit))
                    return set_err(c, LBM_ERR_UNSTABLE, "force record " + std::to_string(first + r) + " (step " +
                                                          std::to_string((long)h[4 * (size_t)r]) + ") is not finite or exceeds force_limit");
            }
    }
    // the state itself, whether or not forces are recorded: the rest population of every live
    // node (a non-finite value reaches it within a step through the moments)
    if (c->initialized) {
        hipError_t e;
        if (ref_mode(c->cfg.bc_mode)) {
            const RefGeo &r = c->rg;
            const int gz = c->D == 3 ? r.g : 0;
            e = count_nonfinite(LBM_FP64, c->rf + r.g + r.rs * ((size_t)r.g + (size_t)r.ly * gz), r.rs, r.rs * r.ly, r.glx,
                                r.gly, r.glz, nullptr, 0, c->dbad, c->stream);
        } else {
            e = count_nonfinite(c->prec, c->A, c->geo.rs, c->geo.rs * c->geo.ny, c->geo.nx, c->geo.ny, c->geo.nz, c->flags,
                                c->geo.nxp, c->dbad, c->stream);
        }
        unsigned bad = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, c->dbad, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return set_err(c, LBM_ERR_HIP, std::string("lbm_sync: ") + hipGetErrorString(e));
        if (bad)
            return set_err(c, LBM_ERR_UNSTABLE, std::to_string(bad) + " nodes hold a non-finite rest population at step " +
                                                    std::to_string(c->step));
    }
    return LBM_OK;
}

extern "C" int lbm_set_step(lbm_ctx *c, long step) {
    if (!c) return LBM_ERR_INVALID;
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_set_step before lbm_initialize");
    if (c->stage != 0) return set_err(c, LBM_ERR_STATE, "lbm_set_step inside a staged step");
    if (step < 0) return set_err(c, LBM_ERR_INVALID, "negative step");
    c->step = step;
    return LBM_OK;
}

// strip the ghost layers of a reference-mode plane (x fastest, interior only)
static void ref_strip(const lbm_ctx *c, const std::vector<double> &pad, double *out) {
    const RefGeo &r = c->rg;
    const int g = r.g, gz = c->D == 3 ? g : 0;
    size_t o = 0;
    for (int k = 0; k < r.glz; k++)
        for (int j = 0; j < r.gly; j++)
            for (int i = 0; i < r.glx; i++, o++)
                out[o] = pad[(size_t)(i + g) + (size_t)r.px * ((size_t)(j + g) + (size_t)r.ly * (k + gz))];
}

extern "C" int lbm_get_fields(lbm_ctx *c, double *p, double *u, double *v, double *w) {
    if (!c) return LBM_ERR_INVALID;
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_get_fields before lbm_initialize");
    double *outs[4] = {p, u, v, w};
    if (c->stage == 0) {
        const int rc = ensure_fields(c);
        if (rc != LBM_OK) return rc;
    }
    if (ref_mode(c->cfg.bc_mode)) {
        std::vector<double> pad(c->rg.plane);
        for (int k = 0; k < 4; k++) {
            if (!outs[k]) continue;
            HIP_TRY(c, hipMemcpyAsync(pad.data(), c->rmac + (size_t)k * c->rg.plane, sizeof(double) * c->rg.plane,
                                      hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            ref_strip(c, pad, outs[k]);
        }
        return LBM_OK;
    }
    double *tmp = nullptr;
    HIP_TRY(c, hipMalloc(&tmp, sizeof(double) * c->N));
    for (int k = 0; k < 4; k++) {
        if (!outs[k]) continue;
        hipError_t e = macros_extract(c->macros, c->flags, tmp, k, c->geo, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(outs[k], tmp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipFree(tmp);
            return set_err(c, LBM_ERR_HIP, std::string("lbm_get_fields: ") + hipGetErrorString(e));
        }
    }
    (void)hipFree(tmp);
    return LBM_OK;
}

extern "C" int lbm_get_populations(lbm_ctx *c, double *f) {
    if (!c || !f) return set_err(c, LBM_ERR_INVALID, "null argument");
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_get_populations before lbm_initialize");
    if (ref_mode(c->cfg.bc_mode)) {
        // between steps the reference's f is the state after Update; inside a staged step, the
        // state after the last stage
        if (c->stage == 0) {
            const int rc = ensure_fields(c);
            if (rc != LBM_OK) return rc;
        }
        const double *src = c->stage == 0 ? c->rfpre : c->rf;
        const RefGeo &r = c->rg;
        std::vector<double> pad(r.plane);   // one direction in the node layout
        for (int i = 0; i < c->Q; i++) {
            // rows of direction i: px elements every rs elements from element i*ds
            HIP_TRY(c, hipMemcpy2DAsync(pad.data(), sizeof(double) * r.px, src + (size_t)i * r.ds, sizeof(double) * r.rs,
                                        sizeof(double) * r.px, (size_t)r.ly * r.lz, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            ref_strip(c, pad, f + (size_t)i * c->N);
        }
        return LBM_OK;
    }
    const FusedArgs a = ctx_fused_args(c);
    double *tmp = nullptr;
    HIP_TRY(c, hipMalloc(&tmp, sizeof(double) * c->N));
    for (int i = 0; i < c->Q; i++) {
        hipError_t e = plain_decode(c->Q, c->prec, c->A, tmp, c->flags, i, a, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(f + (size_t)i * c->N, tmp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipFree(tmp);
            return set_err(c, LBM_ERR_HIP, std::string("lbm_get_populations: ") + hipGetErrorString(e));
        }
    }
    (void)hipFree(tmp);
    return LBM_OK;
}

extern "C" int lbm_set_populations(lbm_ctx *c, const double *f) {
    if (!c || !f) return set_err(c, LBM_ERR_INVALID, "null argument");
    if (!c->initialized) return set_err(c, LBM_ERR_STATE, "lbm_set_populations before lbm_initialize");
    if (ref_mode(c->cfg.bc_mode))
        return set_err(c, LBM_ERR_INVALID, "reference mode keeps ghost state; populations cannot be replaced");
    // the fields stay those of the last step (materialised now, while its input still exists)
    const int rc = ensure_fields(c);
    if (rc != LBM_OK) return rc;
    c->replay_ok = false;
    const FusedArgs a = ctx_fused_args(c);
    double *tmp = nullptr;
    HIP_TRY(c, hipMalloc(&tmp, sizeof(double) * c->N));
    for (int i = 0; i < c->Q; i++) {
        hipError_t e = hipMemcpyAsync(tmp, f + (size_t)i * c->N, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = plain_encode(c->Q, c->prec, c->A, tmp, c->flags, i, a, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipFree(tmp);
            return set_err(c, LBM_ERR_HIP, std::string("lbm_set_populations: ") + hipGetErrorString(e));
        }
    }
    (void)hipFree(tmp);
    return LBM_OK;
}

extern "C" int lbm_get_forces(lbm_ctx *c, double *out, int max, int *count) {
    if (!c) return LBM_ERR_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (count) *count = c->nrec;
    const int n = std::min(max, c->nrec);
    if (out && n > 0) HIP_TRY(c, hipMemcpy(out, c->hist, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost));
    return LBM_OK;
}

extern "C" int lbm_info_get(lbm_ctx *c, lbm_info *info) {
    if (!c || !info) return LBM_ERR_INVALID;
    std::memset(info, 0, sizeof(*info));
    info->lattice = c->Q;
    info->nx = c->geo.nx;
    info->ny = c->geo.ny;
    info->nz = c->geo.nz;
    info->pitch = c->cfg.bc_mode == LBM_BC_PLAIN ? c->geo.nxp : c->rg.px;
    info->nodes = c->N;
    const size_t esz = c->prec == LBM_FP64 ? 8 : c->prec == LBM_FP32 ? 4 : 2;
    info->state_bytes = c->cfg.bc_mode == LBM_BC_PLAIN ? esz * c->Q * c->geo.plane : 8 * c->Q * c->rg.plane;
    info->bytes_per_update = 2.0 * c->Q * (double)esz;
    info->step = c->step;
    info->solid_nodes = c->solid_nodes;
    info->boundary_nodes = c->near_nodes;
    info->nodes_per_lane = c->cfg.bc_mode == LBM_BC_PLAIN ? c->V : 1;
    return LBM_OK;
}
