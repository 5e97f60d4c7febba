// Small helpers shared by the example programs: command-line overrides of
// the run length and a raw dump of a double array for the tests.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

struct RunArgs {
    double maximumTime;
    size_t numFrame;
    std::string dumpFile;  // empty: no dump
};

inline RunArgs parse_run_args(int argc, char** argv, double defTime, size_t defFrames)
{
    RunArgs r{defTime, defFrames, std::string()};
    if (argc > 1) r.maximumTime = std::atof(argv[1]);
    if (argc > 2) r.numFrame = (size_t)std::atol(argv[2]);
    if (argc > 3) r.dumpFile = argv[3];
    return r;
}

// Appends n doubles (interleaved vectors flattened) to path.
inline void dump_append(const std::string& path, const std::vector<double3>& v)
{
    if (path.empty()) return;
    FILE* f = std::fopen(path.c_str(), "ab");
    if (!f) return;
    for (const double3& p : v) {
        const double xyz[3] = {p.x, p.y, p.z};
        std::fwrite(xyz, sizeof(double), 3, f);
    }
    std::fclose(f);
}
