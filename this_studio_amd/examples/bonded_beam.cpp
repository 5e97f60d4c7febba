// An 11-ball bonded cantilever with a 100 kN tip load ramped over 1 s and
// global damping (the scene of the reference's bondedParticleBeam app): the
// load goes through addExternalForceTorque with the externalForceTorque
// launchers, the bonds through addInitialCondition.
// Usage: bonded_beam [maximumTime=5.0] [numFrame=10] [dumpFile]
#include "demsph/DEMSolver.h"
#include "externalForceTorque.h"
#include "common_io.h"

class Beam : public DEMSolver {
public:
    static constexpr size_t kBonds = 10;
    explicit Beam(std::string dump) : DEMSolver(0), dump_(std::move(dump)) {}

    void addExternalForceTorque(const size_t iStep, const double time) override
    {
        (void)iStep;
        const double ramp = time < 1.0 ? time : 1.0;
        std::vector<double3> f(kBonds + 1, make_double3(0.0, 0.0, 0.0));
        f[kBonds] = make_double3(0.0, 0.0, 100.0e3 * ramp);
        load_.setHost(f);
        load_.copyHostToDevice(0);
        DEMBallAccess& b = getBall();
        launchAddConstantForce(b.force(), load_.d_ptr, b.deviceSize(), b.gridDim(), b.blockDim(), 0);
        launchAddGlobalDampingForceTorque(b.force(), b.torque(), b.velocity(), b.angularVelocity(), 0.1,
                                          b.deviceSize(), b.gridDim(), b.blockDim(), 0);
    }
    bool addInitialCondition() override
    {
        addBondedInteraction(getBallPairPointed(), getBallPairPointing(), getBall().positionHostCopy());
        return true;
    }
    void onFrame(const size_t iFrame, const size_t iStep, const double time) override
    {
        const std::vector<double3> p = getBall().positionHostCopy();
        std::printf("frame %zu step %zu t %.6f tip z %.9e\n", iFrame, iStep, time, p.back().z);
        dump_append(dump_, p);
    }

private:
    std::string dump_;
    HostDeviceArray1D<double3> load_;
};

int main(int argc, char** argv)
{
    const RunArgs run = parse_run_args(argc, argv, 5.0, 10);
    const double length = 4.0, r = length / double(Beam::kBonds) / 2.0;
    Beam beam(run.dumpFile);
    beam.setBondedPair(0, 0, 1.0, 200e9, 2.6, 0.0, 0.0, 0.0);
    const double3 zero = make_double3(0.0, 0.0, 0.0);
    beam.addFixedBall(zero, r, 0);
    for (size_t i = 0; i < Beam::kBonds; i++) beam.addBall(make_double3((i + 1) * 2.0 * r, 0.0, 0.0), zero, zero, r, 7800, 0);
    beam.solve(make_double3(0.0, -r, -r), make_double3(length, r, r), zero, 1.e-5, run.maximumTime, run.numFrame,
               "bondedParticleBeam");
    return 0;
}
