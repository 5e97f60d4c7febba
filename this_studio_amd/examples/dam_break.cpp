// A 0.1 x 0.2 x 0.1 m water column at 5 mm spacing collapsing in a
// 0.5 x 0.2 m open tank walled by three layers of dummy particles (the
// scene of the reference's damBreak app), run through the WCSPHSolver
// facade.  Usage: dam_break [maximumTime=5] [numFrame=100] [dumpFile] [stopAfterFrames]
#include "demsph/WCSPHSolver.h"
#include "demsph/pointCloudGeneration.h"
#include "common_io.h"

#include <cmath>

class DamBreak : public WCSPHSolver {
public:
    DamBreak(std::string dump, size_t stopAfter) : WCSPHSolver(0), dump_(std::move(dump)), stopAfter_(stopAfter) {}
    void onFrame(const size_t iFrame, const size_t iStep, const double time) override
    {
        std::printf("frame %zu step %zu t %.6f\n", iFrame, iStep, time);
        dump_append(dump_, getSPH().positionHostCopy());
        if (stopAfter_ && iFrame >= stopAfter_) requestStop();
    }

private:
    std::string dump_;
    size_t stopAfter_;
};

int main(int argc, char** argv)
{
    const RunArgs run = parse_run_args(argc, argv, 5.0, 100);
    const size_t stopAfter = argc > 4 ? (size_t)std::atol(argv[4]) : 0;
    DamBreak tank(run.dumpFile, stopAfter);
    const double H = 0.1;
    double spacing = 0.005;
    const int layers = int(H / spacing);
    if (layers > 0) spacing = H / layers;
    const std::vector<double3> water =
        generateUniformPointCloud(make_double3(0.0, 0.0, 0.0), make_double3(H, 2.0 * H, H), spacing);
    const double3 wall = make_double3(3.0 * spacing, 3.0 * spacing, 3.0 * spacing);
    const double3 lo = make_double3(0.0, 0.0, 0.0) - wall, hi = make_double3(5.0 * H, 2.0 * H, 3.0 * H) + wall;
    const std::vector<double3> shell = generateUniformPointCloud(lo, hi, spacing);
    const double c = 20.0 * std::sqrt(9.81 * H);
    const double3 rest = make_double3(0.0, 0.0, 0.0);
    for (const double3& p : water) tank.addSPH(p, rest, c, spacing, 1000.0, 1.e-3);
    for (const double3& p : shell)
        if (p.x < 0 || p.y < 0 || p.z < 0 || p.x > 5.0 * H || p.y > 2.0 * H) tank.addDummy(p, rest, c, spacing, 1000.0, 1.e-3);
    tank.solve(lo, hi, make_double3(0.0, 0.0, -9.81), 0.25 * 1.3 * spacing / (c + 2.0 * std::sqrt(9.81 * H)),
               run.maximumTime, run.numFrame, "damBreak");
    return 0;
}
