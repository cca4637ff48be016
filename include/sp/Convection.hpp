// Convection.hpp -- drop-in C++ propagator over the C ABI.
//
// Same constructor and methods as the reference class (src/Convection/Convection.h:22-76):
//   Convection(isize, jsize, ksize, dx, nu0, nufreq, cx, cy, cz)
//   DoRK4 / DoEuler (in, out, tStart, dt, timesteps), DoRK4Timestep / DoEulerTimestep.
// Work is enqueued on the library stream; fields are sp::ConvectionField.
#pragma once

#include "ConvectionField.hpp"

namespace sp {

class Convection {
public:
    Convection(int isize, int jsize, int ksize, double dx, double nu0, double nufreq, double cx, double cy,
               double cz)
    {
        check(sp_prop_create(context(), isize, jsize, ksize, dx, nu0, nufreq, cx, cy, cz, &p_));
    }
    ~Convection() { (void)sp_prop_destroy(p_); }
    Convection(const Convection&) = delete;
    Convection& operator=(const Convection&) = delete;

    void DoRK4(ConvectionField& in, ConvectionField& out, double tStart, double dt, int timesteps)
    {
        check(sp_do_rk4(p_, &in.layout(), in.device(), &out.layout(), out.device(), tStart, dt, timesteps));
    }
    void DoEuler(ConvectionField& in, ConvectionField& out, double tStart, double dt, int timesteps)
    {
        check(sp_do_euler(p_, &in.layout(), in.device(), &out.layout(), out.device(), tStart, dt, timesteps));
    }
    void DoRK4Timestep(ConvectionField& in, ConvectionField& out, double t, double dt)
    {
        check(sp_do_rk4_timestep(p_, &in.layout(), in.device(), &out.layout(), out.device(), t, dt));
    }
    void DoEulerTimestep(ConvectionField& in, ConvectionField& out, double t, double dt)
    {
        check(sp_do_euler_timestep(p_, &in.layout(), in.device(), &out.layout(), out.device(), t, dt));
    }

    // extensions: the reference-order kernel path and the Advection2 branch of DoEuler
    void setStaged(bool staged) { check(sp_prop_set_path(p_, staged ? SP_PATH_STAGED : SP_PATH_FUSED)); }
    // times the fused kernel's tiles for `slices` FP64 slices and keeps the fastest
    double autotune(int slices = 1)
    {
        double us = 0.;
        check(sp_prop_autotune(p_, slices, SP_F64, SP_F64, &us));
        return us;
    }
    void setUpwind(bool upwind) { check(sp_prop_set_upwind(p_, upwind ? 1 : 0)); }
    sp_prop* handle() { return p_; }

private:
    sp_prop* p_ = nullptr;
};

}  // namespace sp
