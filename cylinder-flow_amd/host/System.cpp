// System.cpp -- see System.hpp.
#include "System.hpp"

#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>

System::System() { parse("input.in"); }

System::System(const std::string &path) { parse(path); }

// 3D MPI+KOKKOS+MPIIO/System.cpp:5-23
void System::parse(const std::string &path) {
    dim = 3;
    cs2 = 1.0 / 3.0;
    cs = std::sqrt(cs2);
    std::ifstream input(path);
    if (!input) {
        std::cerr << "System: cannot open " << path << std::endl;
        return;
    }
    int Re_i = 0;
    input >> rho0 >> R >> Re_i;
    input >> u0 >> Time >> inter;
    input >> sx >> sy >> sz;
    parsed_ok = static_cast<bool>(input);
    Re = Re_i;
    Ma = u0 / cs;
    miu = rho0 * u0 * R / Re_i;
    tau = u0 * R / Re_i / cs2;
    Time = (int)(std::pow(2 * R, 2) / miu);
    inter = Time / 10;
    // extended "key = value" lines
    std::string line;
    while (std::getline(input, line)) {
        const auto eq = line.find('=');
        if (eq == std::string::npos || line.find_first_not_of('=') == std::string::npos) continue;
        std::string k = line.substr(0, eq), v = line.substr(eq + 1);
        auto trim = [](std::string s) {
            const auto a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
            return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
        };
        k = trim(k);
        v = trim(v);
        if (!k.empty() && !v.empty()) options[k] = v;
    }
    if (options.count("lattice") && options["lattice"] == "D2Q9") {
        dim = 2;
        sz = 1;
    }
}

System::System(int &nx, int &ny) : sx(nx), sy(ny), sz(1), dim(2) {}

// MPI Kokkos/System.cpp:3-18 ; single node: Kokkos/System.cpp:3-18 (R is an int in both)
void System::Initialize() {
    cs2 = 1.0 / 3.0;
    cs = std::sqrt(cs2);
    rho0 = 1.0;
    R = (int)(0.1 * sy);
    Re = 100;
    u0 = 0.1;
    Ma = u0 / cs;
    miu = rho0 * u0 * R / Re;
    Time = single_node ? 100000 : 1000;
    inter = single_node ? 10000 : 100;
    tau = u0 * R / Re / cs2;
    parsed_ok = true;
}

std::string System::option(const std::string &key, const std::string &dflt) const {
    auto it = options.find(key);
    return it == options.end() ? dflt : it->second;
}

// 3D MPI+KOKKOS+MPIIO/System.cpp:25-41 ; MPI Kokkos/System.cpp:20-32
void System::Monitor() {
    if (dim == 3)
        std::cout << "============================" << std::endl << "3D Cylinder Flow" << std::endl;
    else
        std::cout << "2D Cylinder Flow" << std::endl;
    std::cout << "Re    =" << Re << std::endl
              << "Ma    =" << Ma << std::endl
              << "rho   =" << rho0 << std::endl
              << "miu   =" << miu << std::endl
              << "tau   =" << tau << std::endl
              << "Time  =" << Time << std::endl
              << "inter =" << inter << std::endl;
    if (dim == 3)
        std::cout << "nx    =" << sx << std::endl << "ny    =" << sy << std::endl << "nz    =" << sz << std::endl;
    std::cout << "============================" << std::endl;
}
