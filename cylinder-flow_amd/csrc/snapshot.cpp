// snapshot.cpp -- asynchronous field snapshots in the reference's binary Tecplot layout.
//
// Replaces LBM::MPIoutput (3D MPI+KOKKOS+MPIIO/lbm.cpp:419-739; 2D MPI Kokkos/lbm.cpp:
// 459-737).  The reference gathers the fields on the host, runs eight min/max reductions
// and writes through MPI-IO subarray views with the whole solver stopped.  Here the
// fields are copied device -> pinned host memory on the context stream (ordered after the
// step that produced them) and a host thread waits for that copy, computes the min/max and
// writes the file, so the solver can keep stepping.  With one GPU the whole domain is one
// subarray, so the file is the header followed by each variable in FORTRAN (x fastest)
// order, byte for byte what the reference writes from one rank:
//   3D: 332-byte header, then x, y, z, u, v, w, p  (glx*gly*glz doubles each)
//   2D: 276-byte header, then x, y, u, v, p        (glx*gly doubles each)
#include <cstdio>
#include <cstring>

#include "internal.hpp"

using namespace lbm;

struct SnapshotJob {
    std::thread th;
    int rc = LBM_OK;
    std::string err;
};

namespace {

struct Writer {
    std::vector<char> buf;
    void i32(int v) { put(&v, 4); }
    void f32(float v) { put(&v, 4); }
    void f64(double v) { put(&v, 8); }
    void put(const void *p, size_t n) { buf.insert(buf.end(), (const char *)p, (const char *)p + n); }
};

void minmax(const double *a, size_t n, double &lo, double &hi) {
    lo = a[0];
    hi = a[0];
    for (size_t i = 1; i < n; i++) {
        if (a[i] < lo) lo = a[i];
        if (a[i] > hi) hi = a[i];
    }
}

// header of 3D MPI+KOKKOS+MPIIO/lbm.cpp:528-709 / MPI Kokkos/lbm.cpp:548-712
std::vector<char> header(int D, int glx, int gly, int glz, int n, const double *const *f, size_t N) {
    Writer w;
    w.put("#!TDV112", 8);
    w.i32(1);
    w.i32(0);
    w.i32(0);
    const int names3[] = {'x', 'y', 'z', 'u', 'v', 'w', 'p'}, names2[] = {'x', 'y', 'u', 'v', 'p'};
    const int nv = D == 3 ? 7 : 5;
    w.i32(nv);
    for (int k = 0; k < nv; k++) {
        w.i32(D == 3 ? names3[k] : names2[k]);
        w.i32(0);
    }
    w.f32(299.0f);
    for (int ch : {90, 79, 78, 69, 32, 48, 48, 49, 0}) w.i32(ch);   // "ZONE 001"
    if (D == 3) {
        w.i32(-1);
        w.f64((double)n);
        w.i32(0);
    } else {
        w.i32(-1);
        w.i32(-2);
        w.f64((double)n);
    }
    w.i32(-1);
    w.i32(0);
    w.i32(0);
    w.i32(0);
    w.i32(0);
    w.i32(glx);
    w.i32(gly);
    w.i32(D == 3 ? glz : 1);
    w.i32(0);
    w.f32(357.0f);
    w.f32(299.0f);
    for (int k = 0; k < nv; k++) w.i32(2);
    w.i32(0);
    w.i32(0);
    w.i32(-1);
    w.f64(0.0);
    w.f64(4.0);
    w.f64(0.0);
    w.f64(1.0);
    if (D == 3) {
        w.f64(0.0);
        w.f64(1.0);
    }
    const int nf = D == 3 ? 4 : 3;   // u v (w) p
    for (int k = 0; k < nf; k++) {
        double lo, hi;
        minmax(f[k], N, lo, hi);
        w.f64(lo);
        w.f64(hi);
    }
    return w.buf;
}

int write_plt(const char *path, int D, int glx, int gly, int glz, int n, const double *u, const double *v,
              const double *wv, const double *p, std::string &err) {
    const size_t N = (size_t)glx * gly * glz;
    const double *f3[4] = {u, v, wv, p}, *f2[3] = {u, v, p};
    std::vector<char> head = header(D, glx, gly, glz, n, D == 3 ? f3 : f2, N);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        err = std::string("cannot open ") + path;
        return LBM_ERR_INVALID;
    }
    bool ok = std::fwrite(head.data(), 1, head.size(), fp) == head.size();
    std::vector<double> coord(N);
    const int ncoord = D == 3 ? 3 : 2;
    for (int c = 0; c < ncoord && ok; c++) {
        size_t o = 0;
        for (int k = 0; k < glz; k++)
            for (int j = 0; j < gly; j++)
                for (int i = 0; i < glx; i++, o++)
                    coord[o] = c == 0 ? (double)4.0 * i / (glx - 1) : c == 1 ? (double)j / (gly - 1) : (double)k / (glz - 1);
        ok = std::fwrite(coord.data(), sizeof(double), N, fp) == N;
    }
    const double *vars3[4] = {u, v, wv, p}, *vars2[3] = {u, v, p};
    for (int k = 0; k < (D == 3 ? 4 : 3) && ok; k++)
        ok = std::fwrite(D == 3 ? vars3[k] : vars2[k], sizeof(double), N, fp) == N;
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) {
        err = std::string("short write to ") + path;
        return LBM_ERR_INVALID;
    }
    return LBM_OK;
}

}  // namespace

int snapshot_wait(lbm_ctx *c) {
    if (!c->snap) return LBM_OK;
    if (c->snap->th.joinable()) c->snap->th.join();
    const int rc = c->snap->rc;
    if (rc != LBM_OK) c->err = c->snap->err;
    delete c->snap;
    c->snap = nullptr;
    return rc;
}

extern "C" int lbm_snapshot_wait(lbm_ctx *c) {
    if (!c) return LBM_ERR_INVALID;
    return snapshot_wait(c);
}

extern "C" int lbm_snapshot_async(lbm_ctx *c, const char *path, int n) {
    if (!c || !path) return LBM_ERR_INVALID;
    if (!c->initialized) {
        c->err = "lbm_snapshot_async before lbm_initialize";
        return LBM_ERR_STATE;
    }
    int rc = snapshot_wait(c);
    if (rc != LBM_OK) return rc;
    rc = ensure_fields(c);
    if (rc != LBM_OK) return rc;
    const bool ref = c->cfg.bc_mode != LBM_BC_PLAIN;
    const size_t plane = ref ? c->rg.plane : c->N;
    hipEvent_t ev = nullptr;
    auto fail = [&](hipError_t e, const char *what) {
        if (ev) (void)hipEventDestroy(ev);
        c->err = std::string(what) + ": " + hipGetErrorString(e);
        return LBM_ERR_HIP;
    };
    // staging buffers kept in the context (freed by lbm_destroy): the worker thread only reads
    // the pinned copy, and the next snapshot joins it before reusing them
    if (c->snap_elems < 4 * plane) {
        if (c->snap_host) (void)hipHostFree(c->snap_host);
        if (c->snap_dev) (void)hipFree(c->snap_dev);
        c->snap_host = c->snap_dev = nullptr;
        c->snap_elems = 0;
        hipError_t e = hipHostMalloc((void **)&c->snap_host, sizeof(double) * 4 * plane, hipHostMallocDefault);
        if (e != hipSuccess) return fail(e, "hipHostMalloc");
        if (!ref) {
            e = hipMalloc(&c->snap_dev, sizeof(double) * 4 * plane);
            if (e != hipSuccess) return fail(e, "hipMalloc");
        }
        c->snap_elems = 4 * plane;
    }
    double *const pinned = c->snap_host;
    hipError_t e;
    if (ref) {
        e = hipMemcpyAsync(pinned, c->rmac, sizeof(double) * 4 * plane, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) return fail(e, "hipMemcpyAsync");
    } else {
        double *dtmp = c->snap_dev;
        e = hipSuccess;
        for (int k = 0; k < 4 && e == hipSuccess; k++) e = macros_extract(c->macros, c->flags, dtmp + k * plane, k, c->geo, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(pinned, dtmp, sizeof(double) * 4 * plane, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) return fail(e, "snapshot copy");
    }
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ev, c->stream);
    if (e != hipSuccess) return fail(e, "hipEventRecord");

    SnapshotJob *job = new SnapshotJob();
    c->snap = job;
    const int D = c->D, glx = c->geo.nx, gly = c->geo.ny, glz = c->geo.nz;
    const RefGeo rg = c->rg;
    const std::string file(path);
    job->th = std::thread([job, pinned, ev, ref, plane, D, glx, gly, glz, rg, file, n]() {
        hipError_t e2 = hipEventSynchronize(ev);
        (void)hipEventDestroy(ev);
        if (e2 != hipSuccess) {
            job->rc = LBM_ERR_HIP;
            job->err = std::string("snapshot: ") + hipGetErrorString(e2);
            return;
        }
        const size_t N = (size_t)glx * gly * glz;
        std::vector<double> strip;
        const double *fld[4];
        if (ref) {
            strip.resize(4 * N);
            const int g = rg.g, gz = D == 3 ? g : 0;
            for (int k = 0; k < 4; k++) {
                size_t o = 0;
                for (int kk = 0; kk < glz; kk++)
                    for (int j = 0; j < gly; j++)
                        for (int i = 0; i < glx; i++, o++)
                            strip[k * N + o] = pinned[k * plane + (size_t)(i + g) + (size_t)rg.px * ((size_t)(j + g) + (size_t)rg.ly * (kk + gz))];
                fld[k] = strip.data() + k * N;
            }
        } else {
            for (int k = 0; k < 4; k++) fld[k] = pinned + k * plane;
        }
        // macros are stored p, u, v, w
        job->rc = write_plt(file.c_str(), D, glx, gly, glz, n, fld[1], fld[2], fld[3], fld[0], job->err);
    });
    return LBM_OK;
}
