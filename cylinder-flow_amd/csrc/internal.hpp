// internal.hpp -- context and launcher interfaces shared by the kernel files and the
// C ABI (lbm_capi.cpp).  Not installed; include/lbm_capi.h is the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <thread>
#include <vector>

#include "lbm_capi.h"
#include "store.hpp"

namespace lbm {

// Device layout of the plain path.  Node arrays (flags, moments, solid fraction) are x fastest
// with rows padded to `nxp` (a multiple of 64 elements), so every wavefront row segment starts
// aligned:  node (x, y, z) = x + nxp*row,  row = y + ny*z,  plane = nxp*ny*nz.
// A population buffer holds, per row, the Q direction rows one after another (row-interleaved):
//   element (i, x, row) = i*ds + row*rs + x,   ds = nxp, rs = Q*nxp   (Q*plane elements)
// so one row's update reads and writes Q*nxp*sizeof(T) contiguous bytes (plus the source rows'
// stretches) instead of Q strips one plane apart.  ds = plane, rs = nxp gives the layout with
// one plane per direction, kept selectable (LBM_DIRECTION_PLANES) for comparison.
struct Geo {
    int nx, ny, nz, nxp;
    size_t plane;
    size_t ds, rs;
    int xper, yper, zper;
};

// population element of direction i at (x, row)
__host__ __device__ __forceinline__ size_t pel(const Geo &g, int i, uint32_t x, uint32_t row) {
    return (size_t)i * g.ds + (size_t)row * g.rs + x;
}

struct FusedArgs {
    Geo g;
    double tau;      // tau0 + 0.5 (FP64 path: f - (f - feq)/tau, the reference's form)
    double rtau;     // 1/tau rounded to nearest (for div_cr)
    float omega;     // 1/(tau0 + 0.5) (FP32 path)
    StoreParams sp;
};

// node flags
enum : uint8_t { FL_FLUID = 0, FL_SOLID = 1, FL_NEAR = 2 /* fluid with a solid neighbour */ };

// launch shape of every per-node kernel: blocks of `bx` threads along x, `gx` per row
struct Shape {
    int bx, gx;
};
Shape node_shape(const Geo &g);

hipError_t plain_init(int Q, int prec, void *A, void *B, double *macros, const uint8_t *flags, const double *uin,
                      const double *xi, double perturb, uint64_t seed, const FusedArgs &a, hipStream_t s);
// launch shape of the fused kernels: V nodes per lane, at most 256 lanes per block
Shape lane_shape(const Geo &g, int V);
// launch shape of the FP64 kernel k_fused (and of its per-block force slots): 64-lane blocks on
// domains whose population buffer is at most 16 MB (the step is latency-bound there and wants
// every CU), otherwise lane_shape(g, 1)
Shape fp64_shape(const Geo &g, int Q);

// buffers of one fused step
struct StepBufs {
    const void *A;              // populations in (storage type)
    void *B;                    // populations out
    const uint8_t *flags;       // FL_* per node (padded layout)
    const uint32_t *fix;        // FP32/16-bit: nodes the packed path gets wrong (padded indices)
    const uint32_t *fixlk;      // their link words: bit i set when direction i pulls from a solid node
    int nfix;
    const int4 *rowwin;         // FP32/16-bit, per row (null: no mask links anywhere): the lanes
                                // [x, x + y) of the packed kernel hold mask links; their V nodes'
                                // link words start at element z of wlk
    const uint32_t *wlk;        // link words of the window nodes (0: no link, or a fix-list node)
    const uint32_t *link;       // the linked nodes the packed kernel handles (padded indices) ...
    const uint32_t *linklk;     // ... their link words
    const int *link_slot;       // ... and force partial slots (force steps: k_link_force)
    int nlink;
    const void *sfrac;          // solid fraction (FP64: double, else float) or null
    const double *uin;          // inlet velocity per (y, z)
    const int *block_slot;      // force partial slot per block (FP64) or per listed node (-1: none)
    double *partials;
    double *macros;             // p, u, v, w planes (written when macros_on)
    bool dense;                 // FP32/16-bit: one per-node pass over every node (k_dense_v)
                                // instead of the copies, the packed pass and the fix-up
    bool aligned_pull;          // 16-bit, two nodes per lane: aligned loads + lane shifts (k_fast_v)
    unsigned fast_lds;          // dynamic LDS bytes per k_fast_v block (caps its blocks per CU; 0: none)
    hipEvent_t ev_bulk[2];      // when set: recorded just before and after the step's bulk kernel
                                // (k_fast_v, k_dense_v, or k_fused in FP64), on the step's stream
};
hipError_t plain_step(int Q, int prec, int V, const StepBufs &b, bool force, bool macros_on, const FusedArgs &a,
                      hipStream_t s);
hipError_t fast_step_f32(int Q, int V, const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s);
hipError_t fast_step_f16(int Q, int V, const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s);
hipError_t fast_step_bf16(int Q, int V, const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s);
// several D2Q9 FP64 steps per launch (small_kernels.hip): tb_depth() steps at most, T x T tiles
// (tb_tile), tb_tiles(g, T, K) blocks with tb_lds_bytes(T) of LDS each; tb_prepare() before use
// bounce-back links of a node for k_tb2d: bit i-1 when direction i pulls from a solid node
// (after the outlet's clamp), bit 7+i when it pulls across a wall (not for the inlet's unknowns)
std::vector<uint16_t> tb_links(const Geo &g, const std::vector<uint8_t> &flags);
int tb_depth();
size_t tb_lds_bytes(int T);
int tb_tiles(const Geo &g, int T, int K, int *tiles_x);
int tb_tile(const Geo &g, int K, int cus);
hipError_t tb_prepare();
// tb_steps: `steps` (<= K) steps in one launch; the partials of its force records go to record
// slots pend, pend + 1, ... of `partials` (tb_records() slots of ntiles * 3 doubles); *nrec = their
// count.  tb_force_flush sums npend slots into hist[rec0 ...], the first at step step1.
int tb_records(int K);
hipError_t tb_steps(const void *A, void *B, const uint8_t *flags, const uint16_t *lk, const double *uin, double *partials, int pend,
                    int *nrec, const FusedArgs &a, int T, int K, int steps, long step0, int fi, hipStream_t s);
hipError_t tb_force_flush(const double *partials, const Geo &g, int T, int K, double *hist, int rec0, int npend,
                          long step1, int fi, hipStream_t s);
// sum of the nslots force partials into hist[rec] (chunks: force_chunks(nslots) * 3 doubles)
int force_chunks(int nslots);
hipError_t force_reduce(const double *partials, int nslots, double *chunks, double *hist, int rec, long step,
                        hipStream_t s);
// FP64 <-> storage conversion of one direction's populations (physical nodes, x fastest)
hipError_t plain_decode(int Q, int prec, const void *A, double *out, const uint8_t *flags, int dir,
                        const FusedArgs &a, hipStream_t s);
hipError_t plain_encode(int Q, int prec, void *A, const double *in, const uint8_t *flags, int dir,
                        const FusedArgs &a, hipStream_t s);
// number of non-finite values of one direction over a box of nodes (see k_count_nonfinite) into
// *bad (device memory), for storage format prec
hipError_t count_nonfinite(int prec, const void *base, size_t sy, size_t sz, int nx, int ny, int nz,
                           const uint8_t *flags, int fpitch, unsigned *bad, hipStream_t s);
// moment plane k (p, u, v, w) at the physical nodes; 0 at solid nodes
hipError_t macros_extract(const double *macros, const uint8_t *flags, double *out, int k, const Geo &g, hipStream_t s);

// ---- reference mode (FP64, ghost-padded SoA) ----
struct RefGeo {
    int Q, D, g;                 // ghost width 3 (3D) / 2 (2D)
    int glx, gly, glz;           // interior extents
    int lx, ly, lz;              // padded extents
    int px;                      // row pitch in elements (>= lx; 3D rows start 256-byte aligned)
    size_t plane;                // px*ly*lz
    size_t ds, rs;               // population element (ii, i, j, k) = ii*ds + i + rs*(j + ly*k)
    double tau0, u0;
    int variant;                 // D2Q9: 0 the MPI variant, 1 the single-node variant
    unsigned lds;                // dynamic LDS per block of the fused kernel (caps its blocks per CU; 0: none)
};
hipError_t ref_init(const RefGeo &r, double *f, double *ft, double *mac, uint8_t *usr, const int32_t *ran,
                    int body, double cx, double cy, double cz, double r2, uint64_t seed, hipStream_t s);
hipError_t ref_fills(const RefGeo &r, double *f, double *ft, hipStream_t s);
// list (may be null: sweep the interior): the nodes usrbc can change, nlist of them
hipError_t ref_usrbc(const RefGeo &r, double *f, const uint8_t *usr, const uint32_t *list, int nlist, hipStream_t s);
// with_usrbc: the same pass also applies usrbc (in place of a separate ref_usrbc before it)
hipError_t ref_mem(const RefGeo &r, double *f, const uint8_t *usr, double *partials, int nblk,
                   double *hist, int rec, long step, bool with_usrbc, hipStream_t s);
// fpre (may be null): receives the pulled populations, i.e. the reference's f after Update
hipError_t ref_fused(const RefGeo &r, const double *f, double *ft, double *mac, const uint8_t *usr,
                     const double *prof, double *fpre, hipStream_t s);
// the staged step's pull (ft = stream1 of f at the interior nodes) and its Update + next
// Collision in place on f (fpre as for ref_fused)
hipError_t ref_stream(const RefGeo &r, const double *f, double *ft, hipStream_t s);
hipError_t ref_update(const RefGeo &r, double *f, double *mac, const uint8_t *usr, const double *prof, double *fpre,
                      hipStream_t s);
int ref_mem_blocks(const RefGeo &r);

// ---- adjoint (plain mode) ----
struct AdjArgs {
    FusedArgs f;
    double alpha;        // window weight of this step (dJ/dFx)
    double alpha_next;   // window weight of the next step
};
// rowclean: per row (ny*nz bytes), 1 when no node of the row is solid or next to a solid node;
// all_fast: every row is clean and none lies on a periodic y/z edge (the straight-line kernel alone)
hipError_t adj_step(int Q, int prec, const void *H, const void *mu_next, void *mu, const uint8_t *flags,
                    const void *sfrac, const double *uin, void *dJds, double *dJdu_rows, const uint8_t *rowclean,
                    bool all_fast, const AdjArgs &a, hipStream_t s);
hipError_t adj_init(int Q, int prec, const void *mu1, const uint8_t *flags, const double *uin,
                    const double *xi, double perturb, uint64_t seed, double alpha1, double *dJdu_rows,
                    const FusedArgs &a, hipStream_t s);
hipError_t adj_reduce_ds(int prec, const void *dJds, double *out, const Geo &g, hipStream_t s);

}  // namespace lbm

struct SnapshotJob;

struct lbm_ctx {
    lbm_config cfg{};
    int Q = 0, D = 0, prec = 0;
    lbm::Geo geo{};
    size_t N = 0;           // physical nodes
    hipStream_t stream = nullptr;
    std::string err;

    // host-side set-up inputs
    std::vector<uint8_t> solid;     // N
    std::vector<double> sfrac;      // N or empty
    std::vector<double> uin;        // ny*nz
    std::vector<double> xi;         // N or empty (injected perturbation)
    std::vector<int32_t> ran;       // padded box or empty (reference mode)
    bool uin_set = false;
    int solid_nodes = 0, near_nodes = 0;

    // plain-mode device state
    void *A = nullptr, *B = nullptr;
    uint8_t *flags = nullptr;
    uint8_t *rowclean = nullptr;    // per row: 1 when every node is fluid with no solid neighbour
    bool adj_all_fast = false;      // every row clean and off the periodic y/z edges
    void *dsfrac = nullptr;
    double *duin = nullptr;
    double *macros = nullptr;       // 4 planes of FP64
    int *block_slot = nullptr;
    double *partials = nullptr;
    int nslots = 0;
    int tb_k = 0;                   // D2Q9 FP64 small domains: steps per temporally blocked launch (0: off)
    int tb_t = 0;                   // and its tile side
    double *tb_partials = nullptr;  // their per-tile force partials (tb_records(tb_k) record slots)
    uint16_t *tb_lk = nullptr;      // and each node's bounce-back links (tb_links)
    uint32_t *fix = nullptr;        // FP32/16-bit: nodes recomputed per node after the packed path
    uint32_t *fixlk = nullptr;      // their link words
    int nfix = 0;
    int4 *rowwin = nullptr;         // FP32/16-bit: per-row windows of mask links (see StepBufs)
    uint32_t *wlk = nullptr;
    uint32_t *link = nullptr;       // mask-neighbour nodes handled by the packed kernel, their link
    uint32_t *linklk = nullptr;     // words and force slots (for k_link_force)
    int *link_slot = nullptr;
    int nlink = 0;
    int V = 1;                      // nodes per lane of the fused kernel
    bool aligned_pull = true;
    bool dense = false;             // FP32/16-bit: per-node pass over every node (see StepBufs)
    unsigned fast_lds = 0;

    unsigned *dbad = nullptr;       // lbm_sync's count of non-finite populations (device)
    double *noalloc_scratch = nullptr;   // lbm_get_fields_noalloc's staging array on this device
    std::string noalloc_err;

    // force history (device + host copy)
    double *hist = nullptr;
    int hist_cap = 0, nrec = 0, nrec_checked = 0;

    // reference-mode device state
    lbm::RefGeo rg{};
    double *rf = nullptr, *rft = nullptr, *rmac = nullptr, *rprof = nullptr, *rfpre = nullptr;
    uint8_t *rusr = nullptr;
    uint32_t *rsurf = nullptr;      // interior solid nodes with a fluid neighbour (usrbc's work list)
    int nsurf = 0;

    bool initialized = false;
    long step = 0;
    // Fields (and, in reference mode, the populations after Update) are written on request, not
    // every step: fields_fresh says they belong to the current state; otherwise replay_ok says the
    // last step's input buffer is intact, so rerunning that step with the moments on rewrites the
    // same populations and fills them in.
    bool fields_fresh = true;
    bool replay_ok = false;
    // snapshot staging: device buffer (plain mode) and pinned host buffer, kept for reuse
    double *snap_dev = nullptr, *snap_host = nullptr;
    size_t snap_elems = 0;
    // staged step (reference mode): 0 between steps (rf holds the next step's collided
    // populations, rfpre the populations after Update), 1 after lbm_collide, 2 after
    // lbm_boundary, 3 after lbm_stream
    int stage = 0;

    SnapshotJob *snap = nullptr;
};

// materialise the fields of the last step (see lbm_capi.cpp); LBM_OK or an LBM_ERR_* code
int ensure_fields(lbm_ctx *c);
// the address of lbm_get_fields_noalloc's staging array on the current device (noalloc.hip)
int noalloc_setup(double **out, std::string &err);
