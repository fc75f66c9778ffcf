// LBM.cpp -- see LBM.hpp.
#include "LBM.hpp"

#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>

LBM::LBM(int sx, int sy, int sz, double &tau, double &rho0_, double &u0_, const LBMOptions &o)
    : glx(o.single_node ? sx + 1 : sx), gly(o.single_node ? sy + 1 : sy), glz(o.lattice == 9 ? 1 : sz), rho0(rho0_),
      tau0(tau), u0(u0_), opt(o) {
    check(lbm_config_default(&cfg, o.lattice), "lbm_config_default");
    cfg.nx = glx;
    cfg.ny = gly;
    cfg.nz = glz;
    cfg.tau0 = tau0;
    cfg.rho0 = rho0;
    cfg.u0 = u0;
    cfg.bc_mode = o.bc_mode;
    cfg.precision = o.precision;
    cfg.body = o.body;
    if (o.custom_body) {
        cfg.cx = o.cx;
        cfg.cy = o.cy;
        cfg.cz = o.cz;
        cfg.r2 = o.r2;
    } else if (o.lattice == 27) {   // 3D MPI+KOKKOS+MPIIO/lbm.cpp:238
        cfg.cx = cfg.cy = cfg.cz = (double)((glx - 1) / 8);
        const double r = (glx - 1) / 64.0;
        cfg.r2 = r * r;
    } else if (o.single_node) {     // Kokkos/lbm.cpp:115 in padded coordinates (node x = i - 2)
        cfg.cx = 0.25 * sx - 2.0;
        cfg.cy = 0.5 * sy;
        cfg.r2 = 0.01 * sy * sy;
    } else {                        // MPI Kokkos/lbm.cpp:108
        cfg.cx = 0.5 * (glx - 1);
        cfg.cy = 0.5 * (gly - 1);
        const double r = 0.03125 * (glx - 1) / 2.0;
        cfg.r2 = r * r;
    }
    if (o.single_node) {
        cfg.bc_mode = LBM_BC_REFERENCE_SN;
        cfg.force_interval = 0;     // the single-node variant computes no drag
    }
    cfg.inlet = o.inlet;
    cfg.perturb = o.perturb;
    if (o.force_interval >= 0) cfg.force_interval = o.force_interval;
    check(lbm_create(&cfg, &ctx), "lbm_create");
}

LBM::~LBM() {
    if (ctx) lbm_destroy(ctx);
}

void LBM::check(int rc, const char *what) {
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error(ctx));
}

void LBM::Initialize() { check(lbm_initialize(ctx), "lbm_initialize"); }

void LBM::Collision() {
    if (stage != 0) throw std::runtime_error("Collision() called twice without Streaming()/Update()");
    if (opt.staged) check(lbm_collide(ctx), "lbm_collide");
    stage = 1;
}

void LBM::Streaming() {
    if (stage != 1) throw std::runtime_error("Streaming() must follow Collision()");
    if (opt.staged) {
        check(lbm_boundary(ctx), "lbm_boundary");
        check(lbm_stream(ctx), "lbm_stream");
    }
    stage = 2;
}

void LBM::Update() {
    if (stage != 2) throw std::runtime_error("Update() must follow Collision() and Streaming()");
    stage = 0;
    if (opt.staged) check(lbm_update(ctx), "lbm_update");
    else check(lbm_step(ctx, 1), "lbm_step");
    if (opt.log_test_dat && cfg.force_interval > 0) {
        std::vector<double> F = forces();
        const int n = (int)F.size() / 4;
        if (n > logged) {
            // 2*F/(u0^2 D), D = 0.03125*(glx-1) unless set (MPI Kokkos/lbm.cpp:404)
            const double L = opt.cd_length > 0 ? opt.cd_length : 0.03125 * (glx - 1);
            std::ofstream out("test.dat", std::ios::out | std::ios::app);
            out << std::fixed << std::setprecision(17);
            for (int r = logged; r < n; r++)
                out << 2.0 * F[4 * r + 1] / u0 / u0 / L << " " << 2.0 * F[4 * r + 2] / u0 / u0 / L << std::endl;
            logged = n;
        }
    }
}

void LBM::fields(std::vector<double> &p, std::vector<double> &u, std::vector<double> &v, std::vector<double> &w) {
    const size_t N = (size_t)glx * gly * glz;
    p.resize(N);
    u.resize(N);
    v.resize(N);
    w.resize(N);
    check(lbm_get_fields(ctx, p.data(), u.data(), v.data(), w.data()), "lbm_get_fields");
}

std::vector<double> LBM::forces() {
    int n = 0;
    check(lbm_get_forces(ctx, nullptr, 0, &n), "lbm_get_forces");
    std::vector<double> F(4 * (size_t)n);
    if (n) check(lbm_get_forces(ctx, F.data(), n, &n), "lbm_get_forces");
    return F;
}

void LBM::MPIoutput(int n) {
    const std::string name = "output" + std::to_string(n) + ".plt";
    check(lbm_snapshot_async(ctx, name.c_str(), n), "lbm_snapshot_async");
}

// ASCII output of one rank: 3D MPI+KOKKOS+MPIIO/lbm.cpp:741-771 (x y z f_1) and
// MPI Kokkos/lbm.cpp:738-765 (x y u v p)
void LBM::Output(int n) {
    check(lbm_snapshot_wait(ctx), "lbm_snapshot_wait");
    // the single-node 2D variant writes output<n>.dat, x scaled by 4 (Kokkos/lbm.cpp:219-243)
    const std::string name = "output" + std::to_string(n) + (opt.single_node ? ".dat" : "0.dat");
    const double xs = opt.single_node ? 4.0 : 1.0;
    std::ofstream out;
    out << std::setiosflags(std::ios::fixed);
    out.open(name, std::ios::out);
    const size_t N = (size_t)glx * gly * glz;
    if (glz > 1 || cfg.lattice == 27) {
        std::vector<double> f((size_t)cfg.lattice * N);
        check(lbm_get_populations(ctx, f.data()), "lbm_get_populations");
        out << "variables=x,y,z,f" << std::endl;
        out << "zone I=" << glx << ",J=" << gly << ",K=" << glz << std::endl;
        size_t o = 0;
        for (int k = 0; k < glz; k++)
            for (int j = 0; j < gly; j++)
                for (int i = 0; i < glx; i++, o++)
                    out << std::setprecision(8) << std::setiosflags(std::ios::left) << i << " " << j << " " << k << " "
                        << f[N + o] << std::endl;
    } else {
        std::vector<double> p, u, v, w;
        fields(p, u, v, w);
        out << "variables=x,y,u,v,p" << std::endl;
        out << "zone I=" << glx << ",J=" << gly << std::endl;
        size_t o = 0;
        for (int j = 0; j < gly; j++)
            for (int i = 0; i < glx; i++, o++)
                out << std::setprecision(8) << std::setiosflags(std::ios::left) << xs * i / (glx - 1.0) << " "
                    << j / (gly - 1.0) << " " << u[o] << " " << v[o] << " " << p[o] << std::endl;
    }
    std::printf("\nThe result %d is writen\n\n============================\n", n);
}
