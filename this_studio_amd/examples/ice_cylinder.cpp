// A bonded hex sheet of ice balls (r = 0.1 m, 900 kg/m^3) drifting at
// 0.4 m/s onto a triangulated cylinder, with buoyancy and drag applied
// through addExternalForceTorque (the scene of the reference's
// iceCylinderInteraction app), run through the DEMSolver facade.
// Usage: ice_cylinder [maximumTime=50] [numFrame=1000] [dumpFile] [stopAfterFrames]
#include "demsph/DEMSolver.h"
#include "demsph/pointCloudGeneration.h"
#include "externalForceTorque.h"
#include "common_io.h"

#include <cmath>

// Height of the water plane above the centre of a floating sphere: the
// submerged cap fraction (R+h)^2 (2R-h) / (4R^3) equals rhoBall/rhoWater,
// found by bisection on [-R, R].
static double floatingWaterLevel(double R, double rhoBall, double rhoWater)
{
    if (!(R > 0.0) || !(rhoWater > 0.0) || !(rhoBall >= 0.0)) return R;
    const double target = rhoBall / rhoWater;
    if (target <= 0.0) return -R;
    if (target >= 1.0) return R;
    double a = -R, b = R;
    for (int k = 0; k < 80; k++) {
        const double h = 0.5 * (a + b), s = R + h;
        if ((s * s) * (2.0 * R - h) / (4.0 * R * R * R) < target) a = h;
        else b = h;
    }
    return 0.5 * (a + b);
}

class IceSheet : public DEMSolver {
public:
    IceSheet(std::string dump, size_t stopAfter) : DEMSolver(0), dump_(std::move(dump)), stopAfter_(stopAfter) {}
    double waterLevel = 0.0;

    void addExternalForceTorque(const size_t iStep, const double time) override
    {
        (void)iStep;
        (void)time;
        DEMBallAccess& b = getBall();
        launchAddBuoyancyDrag(b.force(), b.position(), b.velocity(), b.radius(), b.inverseMass(), 1000.0, 0.1,
                              make_double3(0.0, 0.0, -9.81), make_double3(-0.4, 0.0, 0.0), waterLevel,
                              b.deviceSize(), b.gridDim(), b.blockDim(), 0);
    }
    bool addInitialCondition() override
    {
        addBondedInteraction(getBallPairPointed(), getBallPairPointing(), getBall().positionHostCopy());
        return true;
    }
    void onFrame(const size_t iFrame, const size_t iStep, const double time) override
    {
        std::printf("frame %zu step %zu t %.6f\n", iFrame, iStep, time);
        dump_append(dump_, getBall().positionHostCopy());
        if (stopAfter_ && iFrame >= stopAfter_) requestStop();
    }

private:
    std::string dump_;
    size_t stopAfter_;
};

int main(int argc, char** argv)
{
    const RunArgs run = parse_run_args(argc, argv, 50.0, 1000);
    const size_t stopAfter = argc > 4 ? (size_t)std::atol(argv[4]) : 0;
    const double tc = 0.0002, r = 0.1, density = 900.0;
    const double mass = r * r * r * pi() * 4.0 / 3.0 * density;
    const double kn = 0.5 * mass * pi() * pi() / tc / tc;
    const double E = kn * 2.0 * r / (r * r * pi());
    IceSheet ice(run.dumpFile, stopAfter);
    ice.waterLevel = floatingWaterLevel(r, density, 1000.0);
    ice.setBondedPair(0, 0, 1.0, E, 2.6, 0.5e6, 0.5e6, 0.1);
    ice.setLinearPair(0, 0, kn, kn / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.1, 0.0, 0.0);
    ice.setLinearPair(0, 1, kn, kn / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.1, 0.0, 0.0);
    std::vector<double3> points;
    HexPBC2D cell;
    generateHexPointCloud(make_double3(0.0, 0.0, 0.0), make_double3(200 * r, 200 * r, 0.0), 2 * r, false, true, points,
                          cell);
    const TriangleMesh cylinder = makeCylinderMesh(make_double3(0.0, 0.0, 0.0), make_double3(0.0, 0.0, 1.0), 1.0, 6.0,
                                                   36, 36);
    const double3 drift = make_double3(-0.4, 0.0, 0.0), zero = make_double3(0.0, 0.0, 0.0);
    for (const double3& p : points) ice.addBall(p, drift, zero, r, density, 0);
    ice.addMeshWall(cylinder.vertices, cylinder.tri0, cylinder.tri1, cylinder.tri2, make_double3(-1.1, 100 * r, 0.0),
                    zero, zero, 1);
    ice.solve(make_double3(-cell.max.x, cell.min.y, -3.0), make_double3(cell.max.x, cell.max.y, 3.0),
              make_double3(0.0, 0.0, -9.81), tc / 50.0, run.maximumTime, run.numFrame, "iceCylinderInteraction");
    return 0;
}
