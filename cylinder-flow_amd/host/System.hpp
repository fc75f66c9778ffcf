// System.hpp -- run parameters, drop-in for the reference's System class.
//
// 3D: System() reads ./input.in (three numeric lines "rho R Re", "u0 Time inter",
// "nx ny nz"; anything after them is ignored by the reference) and derives cs, Ma, miu,
// tau, and overrides Time and inter (3D MPI+KOKKOS+MPIIO/System.cpp:5-23, System.hpp:5-20).
// 2D: System(nx, ny) + Initialize() with the hard-coded values of MPI Kokkos/System.cpp:3-18,
// or of the single-node variant Kokkos/System.cpp:3-18 (Time 100000, inter 10000) when
// single_node is set before Initialize().
//
// Beyond the reference, lines "key = value" after the three numeric lines select this
// solver's options (lattice, bc, precision, body, geometry, inlet, force logging); the
// reference ignores those lines, so one input.in serves both programs.
#ifndef LBM_HOST_SYSTEM_HPP
#define LBM_HOST_SYSTEM_HPP

#include <map>
#include <string>

class System {
public:
    System();                                    // 3D: parse ./input.in
    explicit System(const std::string &path);    // 3D: parse the given file
    System(int &nx, int &ny);                    // 2D
    void Initialize();                           // 2D hard-coded parameters
    void Monitor();

    int sx = 0, sy = 0, sz = 1;
    int Time = 0, inter = 0;
    double miu = 0, u0 = 0, R = 0, rho0 = 0, Ma = 0, tau = 0;
    double Re = 0;   // an int in the 3D reference (System.hpp:18); stored as read
    double cs2 = 0, cs = 0;
    int dim = 3;
    bool single_node = false;   // 2D: the parameters of the single-node variant

    // extended options: key -> value as written in input.in
    std::map<std::string, std::string> options;
    std::string option(const std::string &key, const std::string &dflt) const;
    bool parsed_ok = false;

private:
    void parse(const std::string &path);
};

#endif
