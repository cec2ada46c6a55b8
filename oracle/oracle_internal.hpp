// Internal helpers of the FP64 oracle (test infrastructure only).
#pragma once

#include "../include/gargantua_constants.h"

#include <cmath>
#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

namespace orc {

struct v3 {
    double x, y, z;
};

inline v3 make(const double a[3]) { return {a[0], a[1], a[2]}; }
inline void store(v3 a, double out[3]) {
    out[0] = a.x;
    out[1] = a.y;
    out[2] = a.z;
}
inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 operator*(v3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline v3 operator*(double s, v3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline v3 operator/(v3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(v3 a) { return std::sqrt(dot(a, a)); }
inline v3 normalize(v3 a) { return a / length(a); }

// Split [0, n) into contiguous ranges over `threads` workers.
inline void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t, int)>& body) {
    if (threads <= 1 || n < 2) {
        body(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        int64_t begin = t * chunk, end = std::min<int64_t>(n, begin + chunk);
        if (begin >= end) break;
        pool.emplace_back(body, begin, end, t);
    }
    for (auto& th : pool) th.join();
}

// Dynamic scheduling over [0, n) in blocks, for skewed per-item costs.
void parallel_dynamic(int64_t n, int threads, int64_t block, const std::function<void(int64_t, int64_t)>& body);

}  // namespace orc
