// parareal -- single-GPU driver with the run modes and output of the reference's
// src/Parareal/main.cpp (Compare, Serial, Parallel, Timing), built on the drop-in C++ API
// (include/sp/*.hpp) over libsp_hip.so.  The slices the reference spreads over MPI ranks
// are the batch index of one device (--timeSlices).  Extensions:
// --coarse euler|rk4|rk4spectral|eulerspectral,
// --coarseDtype f64|f32|bf16, --initialFP64 0|1 (f32 / bf16 G: the initial coarse run in FP64),
// --tol <t>, --graphs 0|1, --autotune 0|1.  Energy counters of the Cray nodes are not
// available and not printed.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "sp/Convection.hpp"
#include "sp/ConvectionField.hpp"
#include "sp/ConvectionSolution.hpp"
#include "sp/MatFile.hpp"
#include "sp/Parareal.hpp"
#include "sp/RuntimeConfiguration.hpp"

using namespace sp;

static double wtime()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// fillQ at t with the run's coefficients on the unit cube (ConvectionSolution.h:23-52)
static void fillQ(ConvectionField& q, const RuntimeConfiguration& conf, double t)
{
    sp::fillQ(q, conf.nu0(), conf.nufreq(), conf.cx(), conf.cy(), conf.cz(), t, 0., 1., 0., 1., 0., 1.);
}

// computeErrorReference, src/Parareal/main.cpp:98-139
static double computeErrorReference(const ConvectionField& field, const ConvectionField& reference)
{
    field.SynchronizeHostStorage();
    reference.SynchronizeHostStorage();
    const IJKSize& s = field.calculationDomain();
    double error = 0., maxfield = 0.;
    for (int i = 0; i < s.iSize(); ++i)
        for (int j = 0; j < s.jSize(); ++j)
            for (int k = 0; k < s.kSize(); ++k) {
                const double e = std::abs(field(i, j, k) - reference(i, j, k));
                error = std::max(error, e);
                maxfield = std::max(std::abs(field(i, j, k)), maxfield);
            }
    return error / maxfield;
}

int main(int argc, char** argv)
{
    std::cout << "Initialization...\n" << std::endl;
    RuntimeConfiguration conf(argc, argv);
    PararealOptions opt;
    for (int a = 1; a < argc; ++a) {
        const std::string k = argv[a];
        const std::string v = a + 1 < argc ? argv[a + 1] : "";
        if (k == "--coarse")
            opt.coarse_kind = v == "rk4"             ? SP_COARSE_RK4
                              : v == "rk4spectral"   ? SP_COARSE_RK4_SPECTRAL
                              : v == "eulerspectral" ? SP_COARSE_EULER_SPECTRAL
                                                     : SP_COARSE_EULER;
        if (k == "--coarseDtype") opt.coarse_dtype = v == "f32" ? SP_F32 : v == "bf16" ? SP_BF16 : SP_F64;
        if (k == "--tol") opt.tol = std::atof(v.c_str());
        if (k == "--graphs") opt.use_graphs = v != "0";
        if (k == "--autotune") opt.autotune = v != "0";
        if (k == "--initialFP64") opt.initial_fp64 = v != "0";
    }

    std::cout << "parareal: time-parallel advection-diffusion, all " << conf.timeSlices()
              << " time slices batched on one GPU\n"
              << "  grid " << conf.gridSize() << "^3 (dx = " << conf.dx() << "), T = " << conf.endTime()
              << ", slice length " << conf.timeSliceSize() << "\n"
              << "  nu0 = " << conf.nu0() << ", nufreq = " << conf.nufreq() << ", c = (" << conf.cx() << ", "
              << conf.cy() << ", " << conf.cz() << ")\n"
              << "  fine RK4:  dt = " << conf.dtFine() << ", " << conf.timeStepsFinePerTimeSlice()
              << " steps per slice, CFL " << conf.cflFine() << "\n"
              << "  coarse:    dt = " << conf.dtCoarse() << ", " << conf.timeStepsCoarsePerTimeSlice()
              << " steps per slice, CFL " << conf.cflCoarse() << "\n"
              << "  iterations " << conf.kmax() << ", async fine-step times " << (conf.async() ? "on" : "off")
              << ", mat output " << (conf.mat() ? "on" : "off") << "\n"
              << std::endl;

    try {
        IJKSize domain;
        domain.Init(conf.gridSize(), conf.gridSize(), conf.gridSize());
        KBoundary kboundary;
        kboundary.Init(-2, 2);
        ConvectionField q, qinitial;
        q.Init("q", domain, kboundary);
        qinitial.Init("qinitial", domain, kboundary);
        Convection convection(conf.gridSize(), conf.gridSize(), conf.gridSize(), conf.dx(), conf.nu0(),
                              conf.nufreq(), conf.cx(), conf.cy(), conf.cz());
        // the reference's six-argument form: this process is rank 0 of one, its slice starts at 0
        Parareal<Convection, ConvectionField> parareal(convection, qinitial, q, 0., conf, COMM_WORLD, opt);

        if (conf.mode() == ModeCompare) {
            // cost ratio of the propagators on 4 slices' steps (main.cpp:197-210)
            const int tauSamples = 4;
            check(sp_ctx_sync(context()));
            double tauF = wtime();
            convection.DoRK4(qinitial, qinitial, 0., conf.dtFine(), tauSamples * conf.timeStepsFinePerTimeSlice());
            check(sp_ctx_sync(context()));
            tauF = wtime() - tauF;
            double tauG = wtime();
            convection.DoEuler(qinitial, qinitial, 0., conf.dtCoarse(),
                               tauSamples * conf.timeStepsCoarsePerTimeSlice());
            check(sp_ctx_sync(context()));
            tauG = wtime() - tauG;
            const double tauRatio = tauG / tauF;
            const double Nit_Np = static_cast<double>(conf.kmax()) / conf.timeSlices();
            const double maxSpeedup = 1. / (tauRatio * (1. + Nit_Np) + Nit_Np);

            fillQ(qinitial, conf, 0.);
            qinitial.SynchronizeDeviceStorage();
            double eserial = wtime();
            parareal.DoSerial();
            check(sp_ctx_sync(context()));
            eserial = wtime() - eserial;
            ConvectionField qreference = q;

            fillQ(qinitial, conf, 0.);
            qinitial.SynchronizeDeviceStorage();
            double eparallel = wtime();
            parareal.DoParallel();
            check(sp_ctx_sync(context()));
            eparallel = wtime() - eparallel;

            const double e = computeErrorReference(q, qreference);
            std::cout << "\n"
                      << "Serial run time: " << eserial << "\n"
                      << "Parallel run time: " << eparallel << "\n"
                      << "Speedup: " << eserial / eparallel << "\n"
                      << "Maximal speedup: " << maxSpeedup << "\n"
                      << "Error at end: " << e << "\n"
                      << std::endl;

            MatFile matfile("result.mat");  // main.cpp:251-253
            matfile.addField("q", q);
            matfile.addField("qreference", qreference);
        } else if (conf.mode() == ModeSerial) {
            fillQ(qinitial, conf, 0.);
            qinitial.SynchronizeDeviceStorage();
            double e = wtime();
            parareal.DoSerial();
            check(sp_ctx_sync(context()));
            e = wtime() - e;
            std::cout << "\n" << "Serial run time: " << e << "\n" << std::endl;
        } else if (conf.mode() == ModeParallel) {
            fillQ(qinitial, conf, 0.);
            qinitial.SynchronizeDeviceStorage();
            parareal.DoSerial();
            std::cout << " -- The serial computation is done\n";
            ConvectionField qreference = q;
            double e = wtime();
            parareal.DoParallel();
            check(sp_ctx_sync(context()));
            e = wtime() - e;
            std::cout << " -- The parallel computation is done\n";
            const double error = computeErrorReference(q, qreference);
            std::cout << std::endl;
            std::printf("Parallel run time: %f s\n", e);
            std::printf("Error of parareal: %.4e\n", error);
            std::cout << std::endl;
        } else {
            fillQ(qinitial, conf, 0.);
            qinitial.SynchronizeDeviceStorage();
            std::vector<double> times;
            parareal.DoTimedParallel(times);
            std::cout << "\nTimes:\n";
            for (size_t i = 0; i < times.size(); ++i)
                std::cout << std::scientific << std::setprecision(6) << times[i] << "   " << "\n";
        }
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
