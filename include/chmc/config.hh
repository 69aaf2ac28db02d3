// Host containers with the interfaces of the reference's code-carsten headers:
//   coordinate          (code-carsten/coordinate.hh:6-57)  -- three doubles, [] access, norm, + and -
//   config<T>           (code-carsten/config.hh:15-94)     -- N elements, getN/getd/getSize, [] access,
//                                                             raw binary save/load (load returns 0 / 1)
//   hotstart(config, s) (code-carsten/config.hh:98-105)    -- uniform [0,1)^3 per element from mt19937(seed)
// Positions and momenta handed to libchmc are these AoS double[3] arrays.
#pragma once

#include <cmath>
#include <cstddef>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

class coordinate {
public:
    coordinate() : r{ 0.0, 0.0, 0.0 } {}
    coordinate(double r0, double r1, double r2) : r{ r0, r1, r2 } {}
    explicit coordinate(const double* p) : r{ p[0], p[1], p[2] } {}
    double& operator[](size_t i) { return r[i]; }
    const double& operator[](size_t i) const { return r[i]; }
    double norm() const { return std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]); }
    const double* data() const { return r; }
    double* data() { return r; }

private:
    double r[3];
};
static_assert(sizeof(coordinate) == 3 * sizeof(double), "coordinate must be three packed doubles");

inline coordinate operator+(const coordinate& a, const coordinate& b) { return coordinate(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline coordinate operator-(const coordinate& a, const coordinate& b) { return coordinate(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }

template <class T> class config {
public:
    using value_type = T;
    config(size_t n, size_t d) : N(n), d(d), data_(n) {}
    size_t getN() const { return N; }
    size_t getd() const { return d; }
    size_t getSize() const { return N; }
    size_t storage_size() const { return data_.size() * sizeof(value_type); }
    value_type& operator[](size_t i) { return data_[i]; }
    const value_type& operator[](size_t i) const { return data_[i]; }
    value_type* data() { return data_.data(); }
    const value_type* data() const { return data_.data(); }

    void save(const std::string& path) const
    {
        std::ofstream ofs(path, std::ios::out | std::ios::binary);
        ofs.write(reinterpret_cast<const char*>(data_.data()), storage_size());
    }
    int load(const std::string& path)
    {
        std::ifstream ifs(path, std::ios::in | std::ios::binary);
        if (!ifs) {
            std::cerr << "Error: could not read file from " << path << std::endl;
            return 1;
        }
        ifs.read(reinterpret_cast<char*>(data_.data()), storage_size());
        return 0;
    }

private:
    size_t N, d;
    std::vector<value_type> data_;
};

template <class T> void hotstart(config<T>& c, int seed)
{
    std::mt19937 engine(seed);
    std::uniform_real_distribution<double> dist(0, 1);
    for (size_t i = 0; i < c.getSize(); i++) {
        const double a = dist(engine), b = dist(engine), e = dist(engine);
        c[i] = T(a, b, e);
    }
}
