// A ball dropped onto a two-triangle plane with a linear spring whose
// contact time is 5 ms (the scene of the reference's elasticBounce app),
// run through the DEMSolver facade.  Prints one line per frame.
// Usage: elastic_bounce [maximumTime=2.0] [numFrame=200] [dumpFile]
#include "demsph/DEMSolver.h"
#include "common_io.h"

#include <cmath>

class Bounce : public DEMSolver {
public:
    explicit Bounce(std::string dump) : DEMSolver(0), dump_(std::move(dump)) {}
    void onFrame(const size_t iFrame, const size_t iStep, const double time) override
    {
        const std::vector<double3> p = getBall().positionHostCopy(), v = getBall().velocityHostCopy();
        std::printf("frame %zu step %zu t %.10f z %.10f vz %.10f\n", iFrame, iStep, time, p[0].z, v[0].z);
        dump_append(dump_, {make_double3(time, p[0].z, v[0].z)});
    }

private:
    std::string dump_;
};

int main(int argc, char** argv)
{
    const RunArgs run = parse_run_args(argc, argv, 2.0, 200);
    const double radius = 0.01, contactTime = 0.005, density = 2000.0;
    Bounce scene(run.dumpFile);
    const double3 zero = make_double3(0.0, 0.0, 0.0);
    scene.addBall(zero, zero, zero, radius, density, 0);
    const std::vector<double3> plane = {make_double3(-1.0, -1.0, -1.0), make_double3(1.0, -1.0, -1.0),
                                        make_double3(1.0, 1.0, -1.0), make_double3(-1.0, 1.0, -1.0)};
    scene.addMeshWall(plane, {0, 1}, {1, 2}, {3, 3}, zero, zero, zero, 0);
    const double mass = 4.0 / 3.0 * std::pow(radius, 3.0) * pi() * density;
    const double kn = 0.5 * mass * pi() * pi() / contactTime / contactTime;
    scene.setLinearPair(0, 0, kn, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
    scene.solve(make_double3(-1.0, -1.0, -1.0), make_double3(1.0, 1.0, 1.0), make_double3(0.0, 0.0, -9.81),
                contactTime / 50.0, run.maximumTime, run.numFrame, "elasticBounce");
    return 0;
}
