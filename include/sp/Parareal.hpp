// Parareal.hpp -- drop-in C++ Parareal driver over the C ABI.
//
// Same shape as the reference template (src/Parareal/Parareal.h:38-289):
//   Parareal<PropagatorT, FieldType>(propagator, initial, solution, tStart, conf, comm)
//   DoParallel(), DoTimedParallel(std::vector<double>&), DoSerial()
// The reference runs one slice per MPI rank and takes the rank's start time and the
// communicator (Parareal.h:48-51, called at src/Parareal/main.cpp:193); here all slices are
// the batch index of one GPU, so the one process owns the whole horizon: tStart must be 0
// (rank 0's value, timeSliceSize() * 0) and the communicator is accepted and ignored.  A
// four-argument form (conf, options) needs neither.  The solution field receives the last
// slice's result (the last rank's q).  Extensions select the coarse propagator (DoEuler as
// in the reference, or RK4 in f64 / f32 / bf16 storage), a stopping tolerance and hipGraph
// capture of each iteration.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "Convection.hpp"
#include "RuntimeConfiguration.hpp"

namespace sp {

struct PararealOptions {
    int coarse_kind = SP_COARSE_EULER;
    int coarse_dtype = SP_F64;
    int coarse_substeps = 1;
    double tol = 0.;
    bool use_graphs = false;
    bool autotune = false;
    int correction = -1;  // -1: difference form for f32 / bf16 coarse storage, else the
                          // reference's; 0: the reference's; 1: difference form
    bool initial_fp64 = false;  // f32 / bf16 coarse storage: the initial coarse run in FP64
};

// The communicator argument of the reference constructor: any value converts to it (an
// MPI_Comm handle, MPI_COMM_WORLD, or sp::COMM_WORLD where no MPI header is included).
struct Comm {
    Comm() = default;
    template <typename T>
    Comm(const T&)
    {
    }
};
static const Comm COMM_WORLD{};

template <typename PropagatorT, typename FieldType>
class Parareal {
public:
    // Parareal.h:48-51.  tStart is the calling rank's slice start; with every slice on this
    // GPU the caller is rank 0 of a one-rank job, so only 0 is valid (checked before any
    // device work).
    Parareal(PropagatorT& propagator, FieldType& initial, FieldType& solution, double tStart,
             const RuntimeConfiguration& conf, Comm /*comm*/, const PararealOptions& opt = PararealOptions())
        : Parareal(propagator, initial, solution, start_checked(tStart, conf), opt)
    {
    }

    Parareal(PropagatorT& propagator, FieldType& initial, FieldType& solution, const RuntimeConfiguration& conf,
             const PararealOptions& opt = PararealOptions())
        : initial_(initial), q_(solution), kmax_(conf.kmax())
    {
        sp_parareal_cfg c{};
        c.slices = conf.timeSlices();
        c.kmax = conf.kmax();
        c.end_time = conf.endTime();
        c.steps_fine = conf.timeStepsFine();
        c.steps_coarse = conf.timeStepsCoarse();
        c.coarse_kind = opt.coarse_kind;
        c.coarse_dtype = opt.coarse_dtype;
        c.coarse_substeps = opt.coarse_substeps;
        c.tol = opt.tol;
        c.async_schedule = conf.async() ? 1 : 0;
        c.autotune = opt.autotune ? 1 : 0;
        c.delta_correction = opt.correction < 0 ? (opt.coarse_dtype != SP_F64 ? 1 : 0) : opt.correction;
        c.use_graphs = opt.use_graphs ? 1 : 0;
        c.initial_fp64 = opt.initial_fp64 ? 1 : 0;
        check(sp_parareal_create(propagator.handle(), &c, &pr_));
    }
    ~Parareal() { (void)sp_parareal_destroy(pr_); }
    Parareal(const Parareal&) = delete;
    Parareal& operator=(const Parareal&) = delete;

    void DoParallel()
    {
        check(sp_parareal_run(pr_, &initial_.layout(), initial_.device(), &q_.layout(), q_.device(), &report_));
    }

    // 5*kmax+2 cumulative stamps as in Parareal.h:173-248: after the fine sweep, the
    // receive, the coarse propagation, the update and the send.  Fine and chain times are HIP
    // events on the stream; the chain splits into coarse and update where the corrections
    // were timed (without graphs).  Receive and send take no time on one device.
    void DoTimedParallel(std::vector<double>& times)
    {
        DoParallel();
        times.assign(5 * kmax_ + 2, 0.);
        double t = report_.t_init;
        times[0] = t;
        for (int k = 0; k < kmax_; ++k) {
            if (k < report_.iterations) {
                t += report_.t_fine[k];
                times[5 * k + 1] = times[5 * k + 2] = t;
                t += report_.t_coarse[k] - report_.t_update[k];
                times[5 * k + 3] = t;
                t += report_.t_update[k];
                times[5 * k + 4] = times[5 * k + 5] = t;
            } else {
                for (int q = 1; q <= 5; ++q) times[5 * k + q] = t;
            }
        }
        times[5 * kmax_ + 1] = report_.t_total;
    }

    void DoSerial()
    {
        check(sp_parareal_serial(pr_, &initial_.layout(), initial_.device(), &q_.layout(), q_.device()));
    }

    const sp_parareal_report& report() const { return report_; }

private:
    static const RuntimeConfiguration& start_checked(double tStart, const RuntimeConfiguration& conf)
    {
        if (tStart != 0.)
            throw std::invalid_argument("stencilparareal: Parareal tStart must be 0 (all " +
                                        std::to_string(conf.timeSlices()) +
                                        " time slices run on this device), got " + std::to_string(tStart));
        return conf;
    }

    FieldType& initial_;
    FieldType& q_;
    int kmax_;
    sp_parareal* pr_ = nullptr;
    sp_parareal_report report_{};
};

}  // namespace sp
