// DEMSolver facade over libdemsph: the reference's class, method names,
// argument order and virtual hooks, driving the MI355X DEM scene through the
// C ABI of demsph.h.  Header-only; link with -ldemsph.
//
// Reference: DEMSPH/kernel/DEMSolver.h (setup :1637-1942, getters
// :1541-1626, solve :1961-2049, hooks :1372-1374 and :1536-1539) and the ball
// accessors of DEMSPH/kernel/particle.h:83-260.
//
// Differences a caller can see:
//  * getBall()'s device pointers (force(), position(), ...) are insertion-
//    ordered FP64 views that are valid inside addExternalForceTorque(); the
//    HostCopy() accessors work at any time.
//  * solve() writes the reference's frame files into dir, in its layout:
//    ball, triangleWall (with walls), ballInteraction, bondedInteraction and
//    ballTriangleInteraction (with walls).  An empty dir writes nothing.
//    A subclass may override onFrame() to record results.
//  * Errors throw std::runtime_error (the reference's debug-build behaviour).
#pragma once

#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "demsph_types.h"

class DEMSolver;

// Ball container accessors (particle.h:83-260).
class DEMBallAccess {
public:
    double3* position() { return reinterpret_cast<double3*>(v_.position); }
    double3* velocity() { return reinterpret_cast<double3*>(v_.velocity); }
    double3* angularVelocity() { return reinterpret_cast<double3*>(v_.angularVelocity); }
    double3* force() { return reinterpret_cast<double3*>(v_.force); }
    double3* torque() { return reinterpret_cast<double3*>(v_.torque); }
    double* radius() { return v_.radius; }
    double* inverseMass() { return v_.inverseMass; }

    size_t deviceSize() const { return size(); }
    size_t hostSize() const { return size(); }
    size_t blockDim() const
    {
        const size_t n = size();
        return n == 0 ? 1 : (n < maxThread_ ? n : maxThread_);
    }
    size_t gridDim() const { return (size() + blockDim() - 1) / blockDim(); }

    std::vector<double3> positionHostCopy() { return vec3("position"); }
    std::vector<double3> velocityHostCopy() { return vec3("velocity"); }
    std::vector<double3> angularVelocityHostCopy() { return vec3("angularVelocity"); }
    std::vector<double3> forceHostCopy() { return vec3("force"); }
    std::vector<double3> torqueHostCopy() { return vec3("torque"); }
    std::vector<double> radiusHostCopy() { return scalar("radius"); }
    std::vector<double> inverseMassHostCopy() { return scalar("inverseMass"); }

private:
    friend class DEMSolver;
    demsph_dem* h_ = nullptr;
    demsph_dem_ball_views v_{};
    size_t maxThread_ = 256;

    size_t size() const
    {
        size_t nb = 0;
        if (h_) demsph::check(demsph_dem_sizes(h_, &nb, nullptr, nullptr, nullptr, nullptr, nullptr));
        return nb;
    }
    std::vector<double> scalar(const char* f)
    {
        size_t n = 0;
        demsph::check(demsph_dem_get(h_, f, nullptr, 0, &n));
        std::vector<double> out(n);
        demsph::check(demsph_dem_get(h_, f, out.data(), n, &n));
        return out;
    }
    std::vector<double3> vec3(const char* f)
    {
        const std::vector<double> d = scalar(f);
        std::vector<double3> out(d.size() / 3);
        for (size_t i = 0; i < out.size(); i++) out[i] = make_double3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        return out;
    }
};

class DEMSolver {
public:
    explicit DEMSolver(hipStream_t s) : stream_(s)
    {
        demsph::check(demsph_dem_create(&h_, static_cast<void*>(s)));
        ball_.h_ = h_;
    }
    virtual ~DEMSolver()
    {
        if (h_) (void)demsph_dem_destroy(h_);
    }
    DEMSolver(const DEMSolver&) = delete;
    DEMSolver& operator=(const DEMSolver&) = delete;

    // ------------------------------------------------------------ materials
    void setHertzianPair(const int materialIndexA, const int materialIndexB, const double effectiveYoungsModulus,
                         const double effectiveShearModulus, const double restitutionCoefficient,
                         const double rollingStiffnessToShearStiffnessRatio,
                         const double torsionStiffnessToShearStiffnessRatio, const double slidingFrictionCoefficient,
                         const double rollingFrictionCoefficient, const double torsionFrictionCoefficient)
    {
        demsph::check(demsph_dem_set_hertzian_pair(
            h_, materialIndexA, materialIndexB, effectiveYoungsModulus, effectiveShearModulus, restitutionCoefficient,
            rollingStiffnessToShearStiffnessRatio, torsionStiffnessToShearStiffnessRatio, slidingFrictionCoefficient,
            rollingFrictionCoefficient, torsionFrictionCoefficient));
    }
    void setLinearPair(const int materialIndexA, const int materialIndexB, const double normalStiffness,
                       const double slidingStiffness, const double rollingStiffness, const double torsionStiffness,
                       const double normalDissipation, const double slidingDissipation,
                       const double rollingDissipation, const double torsionDissipation,
                       const double slidingFrictionCoefficient, const double rollingFrictionCoefficient,
                       const double torsionFrictionCoefficient)
    {
        demsph::check(demsph_dem_set_linear_pair(h_, materialIndexA, materialIndexB, normalStiffness,
                                                 slidingStiffness, rollingStiffness, torsionStiffness,
                                                 normalDissipation, slidingDissipation, rollingDissipation,
                                                 torsionDissipation, slidingFrictionCoefficient,
                                                 rollingFrictionCoefficient, torsionFrictionCoefficient));
    }
    void setBondedPair(const int materialIndexA, const int materialIndexB, const double bondRadiusMultiplier,
                       const double bondYoungsModulus, const double normalToShearStiffnessRatio,
                       const double tensileStrength, const double cohesion, const double frictionCoefficient)
    {
        demsph::check(demsph_dem_set_bonded_pair(h_, materialIndexA, materialIndexB, bondRadiusMultiplier,
                                                 bondYoungsModulus, normalToShearStiffnessRatio, tensileStrength,
                                                 cohesion, frictionCoefficient));
    }

    // ------------------------------------------------------------ objects
    void addBall(const double3 position, const double3 velocity, const double3 angularVelocity, const double radius,
                 const double density, const int materialID, const int clumpID = -1)
    {
        demsph::check(demsph_dem_add_balls(h_, 1, &position.x, &velocity.x, &angularVelocity.x, &radius, &density,
                                           &materialID, &clumpID));
    }
    void addFixedBall(const double3 position, const double radius, const int materialID, const int clumpID = -1)
    {
        demsph::check(demsph_dem_add_fixed_balls(h_, 1, &position.x, &radius, &materialID, &clumpID));
    }
    void addClump(std::vector<double3> point, std::vector<double> radius, double3 centroidPosition, double3 velocity,
                  double3 angularVelocity, double mass, symMatrix inertiaTensor, int materialID)
    {
        if (point.size() != radius.size()) return;
        const double I[6] = {inertiaTensor.xx, inertiaTensor.yy, inertiaTensor.zz,
                             inertiaTensor.xy, inertiaTensor.xz, inertiaTensor.yz};
        demsph::check(demsph_dem_add_clump(h_, point.size(), point.empty() ? nullptr : &point[0].x, radius.data(),
                                           &centroidPosition.x, &velocity.x, &angularVelocity.x, mass, I,
                                           materialID));
    }
    void addFixedClump(std::vector<double3> point, std::vector<double> radius, double3 centroidPosition,
                       int materialID)
    {
        if (point.size() != radius.size()) return;
        demsph::check(demsph_dem_add_fixed_clump(h_, point.size(), point.empty() ? nullptr : &point[0].x,
                                                 radius.data(), &centroidPosition.x, materialID));
    }
    void addMeshWall(const std::vector<double3>& vertexPosition, const std::vector<int>& triIndex0,
                     const std::vector<int>& triIndex1, const std::vector<int>& triIndex2, const double3& position,
                     const double3& velocity, const double3& angularVelocity, int materialID)
    {
        if (triIndex1.size() != triIndex0.size() || triIndex2.size() != triIndex0.size()) return;
        demsph::check(demsph_dem_add_mesh_wall(h_, vertexPosition.size(),
                                               vertexPosition.empty() ? nullptr : &vertexPosition[0].x,
                                               triIndex0.size(), triIndex0.data(), triIndex1.data(),
                                               triIndex2.data(), &position.x, &velocity.x, &angularVelocity.x,
                                               materialID));
    }
    void addBondedInteraction(const std::vector<int>& objectPointed, const std::vector<int>& objectPointing,
                              const std::vector<double3>& ballPosition)
    {
        if (objectPointed.size() != objectPointing.size()) return;
        demsph::check(demsph_dem_add_bonded_interactions(h_, objectPointed.size(), objectPointed.data(),
                                                         objectPointing.data(), ballPosition.size(),
                                                         ballPosition.empty() ? nullptr : &ballPosition[0].x));
    }

    // ------------------------------------------------------------ getters
    DEMBallAccess& getBall() { return ball_; }
    std::vector<int> getBallPairPointed() { return ints("objectPointed"); }
    std::vector<int> getBallPairPointing() { return ints("objectPointing"); }
    std::vector<double> getBallPairOverlap() { return doubles("overlap"); }
    std::vector<double3> getBallPairSlidingSpring() { return vec3("slidingSpring"); }
    std::vector<double3> getBallPairRollingSpring() { return vec3("rollingSpring"); }
    std::vector<double3> getBallPairTorsionSpring() { return vec3("torsionSpring"); }
    std::vector<int> getBallTrianglePairPointed() { return ints("triPointed"); }
    std::vector<int> getBallTrianglePairPointing() { return ints("triPointing"); }
    std::vector<double> getBallTrianglePairOverlap() { return doubles("triOverlap"); }
    std::vector<double3> getBallTrianglePairSlidingSpring() { return vec3("triSlidingSpring"); }
    std::vector<double3> getBallTrianglePairRollingSpring() { return vec3("triRollingSpring"); }
    std::vector<double3> getBallTrianglePairTorsionSpring() { return vec3("triTorsionSpring"); }

    // ------------------------------------------------------------ hooks
    // Called twice per step, at t - dt/2 and t, after the contact forces and
    // before the integration (DEMSolver.h:1372-1374, 1376-1405).
    virtual void addExternalForceTorque(const size_t iStep, const double time)
    {
        (void)iStep;
        (void)time;
        defaultHook_ = true;
    }
    // Called once after the first neighbour search and force evaluation; a
    // true return re-initialises the scene (DEMSolver.h:1536-1539, 1994-1997).
    virtual bool addInitialCondition() { return false; }
    // Called at every output frame of solve() (the reference writes VTU
    // files there).  requestStop() inside it ends solve() after the frame.
    virtual void onFrame(const size_t iFrame, const size_t iStep, const double time)
    {
        (void)iFrame;
        (void)iStep;
        (void)time;
    }
    void requestStop() { stop_ = true; }

    // ------------------------------------------------------------ solve
    void solve(const double3 minBoundary, const double3 maxBoundary, const double3 gravity, const double timeStep,
               const double maximumTime, const size_t numFrame, const std::string dir, const size_t deviceID = 0,
               const size_t maximumThread = 256)
    {
        demsph::check_hip(hipSetDevice((int)deviceID), "hipSetDevice");
        if (timeStep <= 0.) {
            std::printf("failed! Time step is less than 0\n");
            return;
        }
        const size_t numStep = size_t(maximumTime / timeStep) + 1;
        size_t frameInterval = numStep;
        if (numFrame > 0) frameInterval = numStep / numFrame;
        if (frameInterval < 1) frameInterval = 1;
        ball_.maxThread_ = maximumThread > 0 ? maximumThread : 256;

        // ball-triangle frames report each contact of the last evaluation
        demsph::check(demsph_dem_record_contacts(h_, dir.empty() ? 0 : 1));
        demsph::check(demsph_dem_initialize(h_, &minBoundary.x, &maxBoundary.x));
        demsph::check(demsph_dem_solve_begin(h_));
        if (addInitialCondition()) demsph::check(demsph_dem_initialize(h_, &minBoundary.x, &maxBoundary.x));
        if (ball_.deviceSize() == 0) {
            std::printf("failed! The number of particles is equal to 0\n");
            return;
        }
        std::printf("Initialization Completed.\n");
        writeFrame(dir, 0, 0, 0.0);

        defaultHook_ = false;
        demsph::check(demsph_dem_set_hook(h_, &DEMSolver::hookTrampoline, this));
        stop_ = false;
        size_t iStep = 0, iFrame = 0;
        while (iStep <= numStep) {
            const size_t next = (iStep / frameInterval + 1) * frameInterval;
            const size_t target = next < numStep + 1 ? next : numStep + 1;
            const int rc = demsph_dem_step(h_, &gravity.x, timeStep, (int)(target - iStep));
            rethrowHookError();
            demsph::check(rc);
            iStep = target;
            if (iStep % frameInterval == 0) {
                size_t s = 0;
                double time = 0.0;
                demsph::check(demsph_dem_clock(h_, &s, &time));
                iFrame++;
                std::printf("Frame %zu at Time %g\n", iFrame, time);
                writeFrame(dir, iFrame, iStep, time);
                onFrame(iFrame, iStep, time);
                if (stop_) break;
            }
        }
        demsph::check(demsph_dem_set_hook(h_, nullptr, nullptr));
        demsph::check(demsph_dem_synchronize(h_));
        demsph::check(demsph_frames_wait());  // ball frames are written in the background
    }

    demsph_dem* handle() { return h_; }

    // outputBallVTU / outputMeshWallVTU / outputBallInteractionVTU /
    // outputBondedInteractionVTU / outputBallTriangleInteractionVTU
    // (DEMSolver.h:908-1370) for one frame.
    void writeFrame(const std::string& dir, const size_t iFrame, const size_t iStep, const double time)
    {
        if (dir.empty()) return;
        demsph::check(demsph_dem_write_vtu_async(h_, dir.c_str(),
                                           DEMSPH_VTU_BALL | DEMSPH_VTU_TRIANGLE_WALL | DEMSPH_VTU_BALL_TRIANGLE |
                                               DEMSPH_VTU_BALL_INTERACTION | DEMSPH_VTU_BONDED_INTERACTION,
                                           (long)iFrame, (long)iStep, time));
    }

private:
    hipStream_t stream_;
    demsph_dem* h_ = nullptr;
    DEMBallAccess ball_;
    bool defaultHook_ = false;
    bool stop_ = false;
    std::exception_ptr hookError_;

    static void hookTrampoline(void* user, size_t iStep, double time)
    {
        DEMSolver* s = static_cast<DEMSolver*>(user);
        if (s->hookError_) return;
        try {
            demsph::check(demsph_dem_views(s->h_, &s->ball_.v_));
            s->addExternalForceTorque(iStep, time);
            // The base implementation ran: no override, so stop scattering
            // views for a hook that adds nothing.
            if (s->defaultHook_) demsph::check(demsph_dem_set_hook(s->h_, nullptr, nullptr));
        } catch (...) {
            s->hookError_ = std::current_exception();
        }
    }
    void rethrowHookError()
    {
        if (hookError_) {
            std::exception_ptr e = hookError_;
            hookError_ = nullptr;
            (void)demsph_dem_set_hook(h_, nullptr, nullptr);
            std::rethrow_exception(e);
        }
    }
    std::vector<int> ints(const char* f)
    {
        size_t n = 0;
        demsph::check(demsph_dem_get(h_, f, nullptr, 0, &n));
        std::vector<int> out(n);
        demsph::check(demsph_dem_get(h_, f, out.data(), n, &n));
        return out;
    }
    std::vector<double> doubles(const char* f)
    {
        size_t n = 0;
        demsph::check(demsph_dem_get(h_, f, nullptr, 0, &n));
        std::vector<double> out(n);
        demsph::check(demsph_dem_get(h_, f, out.data(), n, &n));
        return out;
    }
    std::vector<double3> vec3(const char* f)
    {
        const std::vector<double> d = doubles(f);
        std::vector<double3> out(d.size() / 3);
        for (size_t i = 0; i < out.size(); i++) out[i] = make_double3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        return out;
    }
};
