// main.cpp -- the reference driver loop on the drop-in classes.
//
// 3D (default): 3D MPI+KOKKOS+MPIIO/main.cpp:6-51 -- System from ./input.in, Monitor,
//   LBM, Initialize, MPIoutput(0), then Time steps of Collision/pack/exchange/unpack/
//   Streaming/Update with MPIoutput and the elapsed time every `inter` steps.
// 2D (--2d): MPI Kokkos/main.cpp:6-46 -- nx = 512, ny = 128, hard-coded parameters,
//   Output(0) and Output every `inter` steps, Cd/Cl appended to test.dat every step.
// 2D single node (--2d-single): Kokkos/main.cpp:6-45 -- nx = 512, ny = 128, the single-node
//   variant's parameters (Time 100000, inter 10000), sx+1 by sy+1 nodes, its wall and inlet
//   fill order, no drag, Output(0) and Output every `inter` steps.
// Extra flags (not in the reference): --input PATH, --time N (cap the step count),
//   --check (parse and print the parameters, no GPU work), --report W (3D: print the loop time
//   of the steps after step W, the device drained at both ends, as "loop_seconds S steps N").
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "LBM.hpp"
#include "System.hpp"

static int parse_enum(const std::string &v, std::initializer_list<std::pair<const char *, int>> table, int dflt) {
    for (auto &kv : table)
        if (v == kv.first) return kv.second;
    return dflt;
}

static LBMOptions options_from(const System &s) {
    LBMOptions o;
    o.lattice = s.dim == 2 ? 9 : 27;
    o.bc_mode = parse_enum(s.option("bc", "reference"), {{"reference", LBM_BC_REFERENCE}, {"plain", LBM_BC_PLAIN}}, LBM_BC_REFERENCE);
    o.precision = parse_enum(s.option("precision", "fp64"),
                             {{"fp64", LBM_FP64}, {"fp32", LBM_FP32}, {"fp16s", LBM_FP16S}, {"bf16s", LBM_BF16S}}, LBM_FP64);
    o.body = parse_enum(s.option("body", "sphere"),
                        {{"none", LBM_BODY_NONE}, {"sphere", LBM_BODY_SPHERE}, {"disk", LBM_BODY_SPHERE}, {"zcylinder", LBM_BODY_ZCYL}},
                        LBM_BODY_SPHERE);
    if (s.options.count("radius")) {
        o.custom_body = true;
        o.cx = std::atof(s.option("cx", "0").c_str());
        o.cy = std::atof(s.option("cy", "0").c_str());
        o.cz = std::atof(s.option("cz", "0").c_str());
        const double r = std::atof(s.option("radius", "0").c_str());
        o.r2 = r * r;
    }
    o.inlet = parse_enum(s.option("inlet", "uniform"), {{"uniform", LBM_INLET_UNIFORM}, {"parabolic", LBM_INLET_PARABOLIC}},
                         LBM_INLET_UNIFORM);
    o.perturb = std::atof(s.option("perturb", "0").c_str());
    o.force_interval = std::atoi(s.option("force_interval", "-1").c_str());
    o.cd_length = std::atof(s.option("cd_length", "0").c_str());
    o.log_test_dat = s.option("test_dat", "no") == "yes";
    o.staged = s.option("staged", "no") == "yes";
    return o;
}

int main(int argc, char **argv) {
    std::string input = "input.in";
    long cap = -1, report = -1;
    bool check = false, two_d = false, single = false, staged = false;
    for (int a = 1; a < argc; a++) {
        if (!std::strcmp(argv[a], "--input") && a + 1 < argc) input = argv[++a];
        else if (!std::strcmp(argv[a], "--time") && a + 1 < argc) cap = std::atol(argv[++a]);
        else if (!std::strcmp(argv[a], "--check")) check = true;
        else if (!std::strcmp(argv[a], "--2d")) two_d = true;
        else if (!std::strcmp(argv[a], "--2d-single")) two_d = single = true;
        else if (!std::strcmp(argv[a], "--staged")) staged = true;
        else if (!std::strcmp(argv[a], "--report") && a + 1 < argc) report = std::atol(argv[++a]);
        else {
            std::fprintf(stderr,
                         "usage: %s [--input input.in] [--time N] [--check] [--2d | --2d-single] [--staged] [--report W]\n",
                         argv[0]);
            return 2;
        }
    }
    try {
        if (two_d) {
            int nx = 512, ny = 128;
            System s1(nx, ny);
            s1.single_node = single;
            s1.Initialize();
            s1.Monitor();
            if (check) return 0;
            LBMOptions o;
            o.lattice = 9;
            o.single_node = single;
            o.log_test_dat = !single;
            o.staged = staged;
            double tau = s1.tau, rho0 = s1.rho0, u0 = s1.u0;
            LBM l1(0, s1.sx, s1.sy, 1, tau, rho0, u0, o);
            l1.Initialize();
            l1.Output(0);
            l1.setup_subdomain();
            const long T = cap >= 0 && cap < s1.Time ? cap : s1.Time;
            for (long it = 1; it <= T; it++) {
                l1.Collision();
                l1.pack();
                l1.exchange();
                l1.unpack();
                l1.Streaming();
                l1.Update();
                if (it % s1.inter == 0) l1.Output((int)(it / s1.inter));
            }
            return 0;
        }
        System s1(input);
        if (!s1.parsed_ok) return 1;
        s1.Monitor();
        if (check) return 0;
        LBMOptions o = options_from(s1);
        if (staged) o.staged = true;
        LBM l1(0, s1.sx, s1.sy, s1.sz, s1.tau, s1.rho0, s1.u0, o);
        l1.Initialize();
        l1.MPIoutput(0);
        l1.setup_subdomain();
        const auto start = std::chrono::steady_clock::now();
        const long T = cap >= 0 && cap < s1.Time ? cap : s1.Time;
        auto t_report = start;
        for (long it = 1; it <= T; it++) {
            l1.Collision();
            l1.pack();
            l1.exchange();
            l1.unpack();
            l1.Streaming();
            l1.Update();
            if (s1.inter > 0 && it % s1.inter == 0) {
                l1.MPIoutput((int)(it / s1.inter));
                const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
                std::printf("time=%f\n", t);
            }
            if (it == report) {
                if (lbm_sync(l1.ctx) != LBM_OK) {
                    std::fprintf(stderr, "%s\n", lbm_last_error(l1.ctx));
                    return 1;
                }
                t_report = std::chrono::steady_clock::now();
            }
        }
        const int rc_sync = lbm_sync(l1.ctx);
        const auto t_end = std::chrono::steady_clock::now();   // before joining the snapshot writer
        if (rc_sync != LBM_OK || lbm_snapshot_wait(l1.ctx) != LBM_OK) {
            std::fprintf(stderr, "%s\n", lbm_last_error(l1.ctx));
            return 1;
        }
        if (report >= 0 && report < T)
            std::printf("loop_seconds %.6f steps %ld\n", std::chrono::duration<double>(t_end - t_report).count(),
                        T - report);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
