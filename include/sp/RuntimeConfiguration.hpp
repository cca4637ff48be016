// RuntimeConfiguration.hpp -- the reference's run options (src/Parareal/RuntimeConfiguration.{h,cpp})
// without Boost.program_options: the same option names, defaults, derived quantities and the
// rounding of step totals up to a multiple of the number of time slices.  The reference takes
// the slice count from the MPI world size; here it is the option --timeSlices.
#pragma once

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <iostream>
#include <string>

namespace sp {

enum RunMode { ModeCompare, ModeSerial, ModeParallel, ModeTiming };

class RuntimeConfiguration {
public:
    RuntimeConfiguration(int argc, char** argv)
    {
        for (int a = 1; a < argc; ++a) {
            std::string k = argv[a];
            if (k.rfind("--", 0) != 0) continue;
            std::string v;
            const size_t eq = k.find('=');
            if (eq != std::string::npos) {
                v = k.substr(eq + 1);
                k = k.substr(0, eq);
            } else if (k != "--mat" && k != "--noasync" && k != "--help" && a + 1 < argc) {
                v = argv[++a];
            }
            if (k == "--help") {
                std::cout << "Allowed options: --nu0 --nufreq --cx --cy --cz --gridSize --endTime --timeStepsFine "
                             "--timeStepsCoarse --kmax --mat --noasync --mode --timeSlices\n";
                std::exit(0);
            } else if (k == "--nu0") nu0_ = std::atof(v.c_str());
            else if (k == "--nufreq") nufreq_ = std::atof(v.c_str());
            else if (k == "--cx") cx_ = std::atof(v.c_str());
            else if (k == "--cy") cy_ = std::atof(v.c_str());
            else if (k == "--cz") cz_ = std::atof(v.c_str());
            else if (k == "--gridSize") gridSize_ = std::atoi(v.c_str());
            else if (k == "--endTime") endTime_ = std::atof(v.c_str());
            else if (k == "--timeStepsFine") timeStepsFine_ = std::atoi(v.c_str());
            else if (k == "--timeStepsCoarse") timeStepsCoarse_ = std::atoi(v.c_str());
            else if (k == "--kmax") kmax_ = std::atoi(v.c_str());
            else if (k == "--timeSlices") timeSlices_ = std::atoi(v.c_str());
            else if (k == "--mat") mat_ = true;
            else if (k == "--noasync") async_ = false;
            else if (k == "--mode") {
                std::string opt = v;
                std::transform(opt.begin(), opt.end(), opt.begin(), ::tolower);
                if (opt == "compare") mode_ = ModeCompare;
                else if (opt == "parallel") mode_ = ModeParallel;
                else if (opt == "serial") mode_ = ModeSerial;
                else if (opt == "timing") mode_ = ModeTiming;
                else {
                    std::cerr << "Mode " << v << " not recognized.\n" << "Defaulting to Compare\n";
                    mode_ = ModeCompare;
                }
            }
        }
        // RuntimeConfiguration.cpp:133-151
        if (timeStepsFine_ % timeSlices_ != 0) {
            timeStepsFine_ = (timeStepsFine_ / timeSlices_ + 1) * timeSlices_;
            std::cout << "Warning: timeStepsFine is not a multiple of timeSlices\n"
                      << " -- Changing timeStepsFine to " << timeStepsFine_ << "\n";
        }
        if (timeStepsCoarse_ % timeSlices_ != 0) {
            timeStepsCoarse_ = (timeStepsCoarse_ / timeSlices_ + 1) * timeSlices_;
            std::cout << "Warning: timeStepsCoarse is not a multiple of timeSlices\n"
                      << " -- Changing timeStepsCoarse to " << timeStepsCoarse_ << "\n\n";
        }
    }

    double nu0() const { return nu0_; }
    double nufreq() const { return nufreq_; }
    double cx() const { return cx_; }
    double cy() const { return cy_; }
    double cz() const { return cz_; }
    int timeSlices() const { return timeSlices_; }
    int timeStepsFine() const { return timeStepsFine_; }
    int timeStepsCoarse() const { return timeStepsCoarse_; }
    int gridSize() const { return gridSize_; }
    double endTime() const { return endTime_; }
    int kmax() const { return kmax_; }
    bool mat() const { return mat_; }
    bool async() const { return async_; }
    RunMode mode() const { return mode_; }

    // RuntimeConfiguration.h:54-111
    double dx() const { return 1. / gridSize_; }
    double dtFine() const { return endTime_ / timeStepsFine_; }
    double dtCoarse() const { return endTime_ / timeStepsCoarse_; }
    int timeStepsFinePerTimeSlice() const { return timeStepsFine_ / timeSlices_; }
    int timeStepsCoarsePerTimeSlice() const { return timeStepsCoarse_ / timeSlices_; }
    double timeSliceSize() const { return endTime_ / timeSlices_; }
    double cflFine() const { return cfl(dtFine()); }
    double cflCoarse() const { return cfl(dtCoarse()); }

private:
    double cfl(double dt) const
    {
        const double c = std::max(std::max(cx_, cy_), cz_);
        const double dt_dx = dt / dx();
        const double dt_dx2 = dt_dx / dx();
        return std::max(3. / 2. * nu0_ * dt_dx2, c * dt_dx);
    }

    // defaults: RuntimeConfiguration.cpp:7-12
    double nu0_ = 1., nufreq_ = 0., cx_ = 1., cy_ = 1., cz_ = 1.;
    int gridSize_ = 32;
    double endTime_ = 0.05;
    int timeSlices_ = 1;
    int timeStepsFine_ = 128, timeStepsCoarse_ = 32;
    int kmax_ = 5;
    bool mat_ = false, async_ = true;
    RunMode mode_ = ModeCompare;
};

}  // namespace sp
