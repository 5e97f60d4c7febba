// WCSPHSolver facade over libdemsph: the reference's class, method names and
// argument order, driving the MI355X WCSPH scene through the C ABI of
// demsph.h.  Header-only; link with -ldemsph.
//
// Reference: DEMSPH/kernel/WCSPHSolver.h (addSPH/addDummy :523-541,
// getSPH/getDummy :562-564, solve :567-624) and the WCSPH container
// accessors of DEMSPH/kernel/particle.h:829-835.
//
// solve() writes the fluid and dummy frame files into dir as the reference
// does (SPH_FFFF.vtu, dummy_FFFF.vtu, same layout; an empty dir writes
// nothing).  A subclass may override onFrame() to record results.  Errors
// throw std::runtime_error.
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "demsph_types.h"

class WCSPHSolver;

// Fluid or dummy container accessors (particle.h:829-835), insertion order.
class WCSPHAccess {
public:
    size_t hostSize() const { return dummy_ ? uploaded_.size() + hd_.size() : hf_.size(); }
    size_t deviceSize() const
    {
        size_t nf = 0, nd = 0;
        demsph::check(demsph_sph_sizes(h_, &nf, &nd, nullptr));
        return dummy_ ? nd : nf;
    }
    std::vector<double3> positionHostCopy()
    {
        return dummy_ ? staged(true) : vec3("position");
    }
    std::vector<double3> velocityHostCopy() { return dummy_ ? staged(false) : vec3("velocity"); }
    std::vector<double3> normalHostCopy() { return dummy_ ? vec3("dummyNormal") : std::vector<double3>(); }
    std::vector<double> densityHostCopy() { return scalar("density"); }
    std::vector<double> pressureHostCopy() { return scalar("pressure"); }
    std::vector<double> densityChangeHostCopy() { return scalar("densityChange"); }

private:
    friend class WCSPHSolver;
    struct Row {
        double3 p, v;
        double c, dx, rho0, nu;
    };
    demsph_sph* h_ = nullptr;
    bool dummy_ = false;
    std::vector<Row> hf_, hd_;

    std::vector<Row> uploaded_;  // dummies already handed to the library

    std::vector<double3> staged(bool pos) const
    {
        std::vector<double3> out;
        for (const Row& r : uploaded_) out.push_back(pos ? r.p : r.v);
        for (const Row& r : hd_) out.push_back(pos ? r.p : r.v);
        return out;
    }
    std::vector<double> scalar(const char* f)
    {
        if (dummy_) return std::vector<double>();
        size_t n = 0;
        demsph::check(demsph_sph_get(h_, f, nullptr, 0, &n));
        std::vector<double> out(n);
        demsph::check(demsph_sph_get(h_, f, out.data(), n, &n));
        return out;
    }
    std::vector<double3> vec3(const char* f)
    {
        size_t n = 0;
        demsph::check(demsph_sph_get(h_, f, nullptr, 0, &n));
        std::vector<double> d(n);
        demsph::check(demsph_sph_get(h_, f, d.data(), n, &n));
        std::vector<double3> out(n / 3);
        for (size_t i = 0; i < out.size(); i++) out[i] = make_double3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        return out;
    }
};

class WCSPHSolver {
public:
    explicit WCSPHSolver(hipStream_t s) : stream_(s)
    {
        demsph::check(demsph_sph_create(&h_, static_cast<void*>(s)));
        fluid_.h_ = h_;
        dummy_.h_ = h_;
        dummy_.dummy_ = true;
    }
    virtual ~WCSPHSolver()
    {
        if (h_) (void)demsph_sph_destroy(h_);
    }
    WCSPHSolver(const WCSPHSolver&) = delete;
    WCSPHSolver& operator=(const WCSPHSolver&) = delete;

    // Staged on the host and uploaded by solve(); entries with density <= 0
    // or sound speed <= 0 are skipped (WCSPHSolver.h:523-541).
    void addSPH(double3 position, double3 velocity, double soundSpeed, double spacing, double density,
                double viscosity)
    {
        if (density <= 0. || soundSpeed <= 0.) return;
        fluid_.hf_.push_back({position, velocity, soundSpeed, spacing, density, viscosity});
    }
    void addDummy(double3 position, double3 velocity, double soundSpeed, double spacing, double density,
                  double viscosity)
    {
        if (density <= 0. || soundSpeed <= 0.) return;
        dummy_.hd_.push_back({position, velocity, soundSpeed, spacing, density, viscosity});
    }

    WCSPHAccess& getSPH() { return fluid_; }
    WCSPHAccess& getDummy() { return dummy_; }

    // Equation of state: 0 linear (the reference), 1 Tait with exponent gamma.
    void setEquationOfState(int eos, double gamma = 7.0) { demsph::check(demsph_sph_set_eos(h_, eos, gamma)); }

    virtual void onFrame(const size_t iFrame, const size_t iStep, const double time)
    {
        (void)iFrame;
        (void)iStep;
        (void)time;
    }
    void requestStop() { stop_ = true; }

    void solve(const double3 minBoundary, const double3 maxBoundary, const double3 gravity, const double timeStep,
               const double maximumTime, const size_t numFrame, const std::string dir, const size_t deviceID = 0,
               const size_t maximumThread = 256)
    {
        (void)maximumThread;
        demsph::check_hip(hipSetDevice((int)deviceID), "hipSetDevice");
        if (timeStep <= 0.) {
            std::printf("failed! Time step is less than 0\n");
            return;
        }
        const size_t numStep = size_t(maximumTime / timeStep) + 1;
        size_t frameInterval = numStep;
        if (numFrame > 0) frameInterval = numStep / numFrame;
        if (frameInterval < 1) frameInterval = 1;

        upload(fluid_.hf_, false);
        upload(dummy_.hd_, true);
        fluid_.hf_.clear();
        dummy_.uploaded_.insert(dummy_.uploaded_.end(), dummy_.hd_.begin(), dummy_.hd_.end());
        dummy_.hd_.clear();
        demsph::check(demsph_sph_initialize(h_, &minBoundary.x, &maxBoundary.x, &gravity.x));
        std::printf("Initialization Completed.\n");
        writeFrame(dir, 0, 0, 0.0);

        stop_ = false;
        size_t iStep = 0, iFrame = 0;
        double time = 0.0;
        while (iStep <= numStep) {
            const size_t next = (iStep / frameInterval + 1) * frameInterval;
            const size_t target = next < numStep + 1 ? next : numStep + 1;
            demsph::check(demsph_sph_step(h_, timeStep, (int)(target - iStep)));
            for (size_t k = iStep; k < target; k++) time += timeStep;  // the reference's running sum
            iStep = target;
            if (iStep % frameInterval == 0) {
                iFrame++;
                std::printf("Frame %zu at Time %g\n", iFrame, time);
                writeFrame(dir, iFrame, iStep, time);
                onFrame(iFrame, iStep, time);
                if (stop_) break;
            }
        }
        demsph::check(demsph_sph_synchronize(h_));
        demsph::check(demsph_frames_wait());  // fluid frames are written in the background
    }

    demsph_sph* handle() { return h_; }

    // outputWCSPHVTU / outputDummyVTU (WCSPHSolver.h:360-520) for one frame.
    void writeFrame(const std::string& dir, const size_t iFrame, const size_t iStep, const double time)
    {
        if (dir.empty()) return;
        demsph::check(demsph_sph_write_vtu_async(h_, dir.c_str(), DEMSPH_VTU_SPH | DEMSPH_VTU_DUMMY, (long)iFrame, (long)iStep,
                                           time));
    }

private:
    hipStream_t stream_;
    demsph_sph* h_ = nullptr;
    WCSPHAccess fluid_, dummy_;
    bool stop_ = false;

    void upload(const std::vector<WCSPHAccess::Row>& rows, bool dummy)
    {
        const size_t n = rows.size();
        if (n == 0) return;
        std::vector<double> p(3 * n), v(3 * n), c(n), dx(n), r(n), nu(n);
        for (size_t i = 0; i < n; i++) {
            p[3 * i] = rows[i].p.x; p[3 * i + 1] = rows[i].p.y; p[3 * i + 2] = rows[i].p.z;
            v[3 * i] = rows[i].v.x; v[3 * i + 1] = rows[i].v.y; v[3 * i + 2] = rows[i].v.z;
            c[i] = rows[i].c; dx[i] = rows[i].dx; r[i] = rows[i].rho0; nu[i] = rows[i].nu;
        }
        if (dummy)
            demsph::check(demsph_sph_add_dummy(h_, n, p.data(), v.data(), c.data(), dx.data(), r.data(), nu.data()));
        else
            demsph::check(demsph_sph_add_fluid(h_, n, p.data(), v.data(), c.data(), dx.data(), r.data(), nu.data()));
    }
};
