// LBM.hpp -- drop-in for the reference's LBM struct (3D MPI+KOKKOS+MPIIO/lbm.hpp:106-233,
// MPI Kokkos/lbm.hpp:64-128) over the C ABI of liblbm_hip.so.
//
// The member functions keep their names, call order and meaning:
//   Initialize()                 LBM::Initialize   -> lbm_initialize
//   Collision(); Streaming();    one time step: on the GPU collide, stream, boundaries and
//   Update()                     the moment update are one fused kernel, launched by Update()
//                                after Collision() and Streaming() were called (in that order)
//   setup_subdomain(), pack(),   no-ops: one GPU holds the whole domain and ghost handling
//   exchange(), unpack()         happens inside the kernel
//   Boundary()                   no-op (declared but never defined in the reference)
//   MPIoutput(n)                 binary Tecplot "output<n>.plt", written asynchronously
//   Output(n)                    ASCII "output<n><rank>.dat" as the reference writes it
//                                (single-node 2D variant: "output<n>.dat", Kokkos/lbm.cpp:219-243)
// Device data are owned by the context; fields()/populations() copy them to the host (the
// reference exposes UVM views instead).  Cd/Cl are appended to test.dat like the 2D MPI
// variant (MPI Kokkos/lbm.cpp:398-407) when force logging is on.
#ifndef LBM_HOST_LBM_HPP
#define LBM_HOST_LBM_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_capi.h"

struct LBMOptions {
    int lattice = 27;
    int bc_mode = LBM_BC_REFERENCE;
    int precision = LBM_FP64;
    int body = LBM_BODY_SPHERE;
    bool custom_body = false;
    double cx = 0, cy = 0, cz = 0, r2 = 0;
    int inlet = LBM_INLET_UNIFORM;
    double perturb = 0.0;
    int force_interval = -1;      // -1: the reference's default (2D every step, 3D none)
    double cd_length = 0.0;       // Cd = 2F/(u0^2 * cd_length); 0: the reference's 0.03125*(glx-1)
    bool log_test_dat = false;
    bool single_node = false;     // 2D single-node variant: LBM(sx, sy, ...) holds sx+1 by sy+1
                                  // nodes (Kokkos/lbm.hpp:26), its disk and boundary rules
    bool staged = false;          // reference modes: Collision(), Streaming() and Update() each run
                                  // their own stage (lbm_collide, lbm_boundary + lbm_stream,
                                  // lbm_update) instead of one fused step in Update()
};

struct LBM {
    // LBM(comm, sx, sy, sz, tau, rho0, u0): the communicator is accepted and ignored
    template <class Comm>
    LBM(Comm, int sx, int sy, int sz, double &tau, double &rho0, double &u0, const LBMOptions &o = LBMOptions())
        : LBM(sx, sy, sz, tau, rho0, u0, o) {}
    LBM(int sx, int sy, int sz, double &tau, double &rho0, double &u0, const LBMOptions &o = LBMOptions());
    ~LBM();
    LBM(const LBM &) = delete;
    LBM &operator=(const LBM &) = delete;

    void Initialize();
    void Collision();
    void setup_subdomain() {}
    void pack() {}
    void exchange() {}
    void unpack() {}
    void Streaming();
    void Boundary() {}
    void Update();
    void MPIoutput(int n);
    void Output(int n);

    // host copies of the device state (x fastest)
    void fields(std::vector<double> &p, std::vector<double> &u, std::vector<double> &v, std::vector<double> &w);
    std::vector<double> forces();

    int glx, gly, glz;
    double rho0, tau0, u0;
    lbm_ctx *ctx = nullptr;
    lbm_config cfg{};
    LBMOptions opt;

private:
    void check(int rc, const char *what);
    int stage = 0;       // 1 after Collision, 2 after Streaming
    int logged = 0;      // force records already written to test.dat
};

#endif
