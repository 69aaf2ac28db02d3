// The xyz trajectory files of the reference (the on-disk interchange format): print_xyz
// (identical_particles.cpp:261-266), read_xyz of the last configuration (:89-146), how_many_confs_xyz
// and read_next_confs_xyz (:148-194).  One configuration is
//     "     N"
//     "trajectory= T  kinetic_energy= K  potential= V"
//     N lines "name  x y z" (%-20.12g)
// Differences, both deliberate: read_xyz_last returns the trajectory number of the last
// configuration to the caller, which stores it as istart (the reference writes it into a by-value
// copy and loses it, Q14), and parse errors are exceptions instead of Kokkos::abort.
#pragma once

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace chmc_host {

inline void print_xyz(FILE* f, const std::string& name, int traj, double K, double V, const std::vector<double>& xyz)
{
    const int N = (int)(xyz.size() / 3);
    std::fprintf(f, "     %d\n", N);
    std::fprintf(f, "trajectory= %d  kinetic_energy= %.12g  potential= %.12g\n", traj, K, V);
    for (int i = 0; i < N; i++)
        std::fprintf(f, "%s  %-20.12g %-20.12g %-20.12g\n", name.c_str(), xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    std::fflush(f);
}

// number of configurations of N atoms in an open file (rewound afterwards)
inline int how_many_confs_xyz(FILE* f, int N)
{
    int lines = 0, ch;
    while ((ch = std::fgetc(f)) != EOF)
        if (ch == '\n') lines++;
    std::rewind(f);
    if (lines % (N + 2) != 0)
        throw std::runtime_error("xyz file contains " + std::to_string(lines) + " lines, not a multiple of N+2=" +
                                 std::to_string(N + 2));
    return lines / (N + 2);
}

// the next configuration from the current position of an open file; returns its trajectory number
inline int read_next_confs_xyz(FILE* f, int N, const std::string& name, std::vector<double>& xyz)
{
    char line[4096];
    if (!std::fgets(line, sizeof(line), f)) throw std::runtime_error("xyz: unexpected end of file");
    int n = 0;
    if (std::sscanf(line, " %d", &n) != 1 || n != N) throw std::runtime_error("xyz: atom count line does not hold N");
    if (!std::fgets(line, sizeof(line), f)) throw std::runtime_error("xyz: unexpected end of file");
    int traj = 0;
    if (std::sscanf(line, " trajectory= %d", &traj) != 1) throw std::runtime_error("xyz: missing 'trajectory=' header");
    xyz.resize(3 * (size_t)N);
    char id[1024];
    for (int i = 0; i < N; i++) {
        if (!std::fgets(line, sizeof(line), f)) throw std::runtime_error("xyz: unexpected end of file");
        if (std::sscanf(line, "%1023s %lf %lf %lf", id, &xyz[3 * i], &xyz[3 * i + 1], &xyz[3 * i + 2]) != 4)
            throw std::runtime_error("error in reading the file");
        if (name != id)
            throw std::runtime_error(std::string("name in the xyz file: ") + id + " does not match the name in the input file: " + name);
    }
    return traj;
}

// the last configuration of a file and its trajectory number
inline int read_xyz_last(const std::string& path, int N, const std::string& name, std::vector<double>& xyz)
{
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw std::runtime_error("error in opening file " + path);
    int traj = 0;
    try {
        const int confs = how_many_confs_xyz(f, N);
        if (confs < 1) throw std::runtime_error("xyz file " + path + " holds no configuration");
        for (int c = 0; c < confs; c++) traj = read_next_confs_xyz(f, N, name, xyz);
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    return traj;
}

}  // namespace chmc_host
