// convection -- single-GPU accuracy driver with the options and output of the reference's
// src/Convection/main.cpp:53-409: advance the exact solution's initial state with RK4
// (--order 4, default) or forward Euler (--order 1) and report the relative L-infinity
// error against exactQ at the end time.  Writes init.mat (the initial field) and, with
// --mat, convection_0_0_0.mat (a cell holding the initial state) and result_0_0_0.mat
// (initial field, result, error field, exact field), as the reference does on one process.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "sp/Convection.hpp"
#include "sp/ConvectionField.hpp"
#include "sp/ConvectionSolution.hpp"
#include "sp/MatFile.hpp"

using namespace sp;

namespace {

struct HeatConfiguration {  // src/Convection/main.cpp:53-65 and defaults at :70-83
    double nu0 = 1., nufreq = 0., cx = 1., cy = 1., cz = 1.;
    int gridsize = 32, timesteps = 512;
    double endtime = 0.05, cfl = 0., dx = 1. / 32, dt = 0.05 / 512;
    bool mat = false, rk = true;
};

[[noreturn]] void usage(int rc)
{
    std::cout << "Allowed options:\n"
                 "  --help                Produce this help message\n"
                 "  --nu0 arg             Initial diffusion coefficient\n"
                 "  --nufreq arg          Frequency of diffusion coefficient\n"
                 "  --cx arg              Advection velocity in x direction\n"
                 "  --cy arg              Advection velocity in y direction\n"
                 "  --cz arg              Advection velocity in z direction\n"
                 "  --gridSize arg        Number of gridpoints along each direction\n"
                 "  --timeSteps arg       Number of timesteps\n"
                 "  --endTime arg         Time to simulate\n"
                 "  --cfl arg             Fixed CFL condition\n"
                 "  --order arg           Order of timestepping (valid values are 1 and 4)\n"
                 "  --mat                 Output intermediate states\n";
    std::exit(rc);
}

// Option semantics of parseCommandLine (src/Convection/main.cpp:67-213), including its
// order of evaluation: endTime, gridSize, timeSteps, then cfl (which fixes the step count
// when the grid is given, else the grid from dt = endTime / timeSteps).
HeatConfiguration parse(int argc, char** argv)
{
    HeatConfiguration c;
    bool has_end = false, has_grid = false, has_steps = false, has_cfl = false;
    double end = 0., cfl = 0.;
    int grid = 0, steps = 0, order = 4;
    for (int a = 1; a < argc; ++a) {
        std::string k = argv[a], v;
        if (k.rfind("--", 0) != 0) continue;
        const size_t eq = k.find('=');
        if (eq != std::string::npos) {
            v = k.substr(eq + 1);
            k = k.substr(0, eq);
        } else if (k != "--help" && k != "--mat" && a + 1 < argc) {
            v = argv[++a];
        }
        if (k == "--help") usage(0);
        else if (k == "--mat") c.mat = true;
        else if (k == "--nu0") c.nu0 = std::atof(v.c_str());
        else if (k == "--nufreq") c.nufreq = std::atof(v.c_str());
        else if (k == "--cx") c.cx = std::atof(v.c_str());
        else if (k == "--cy") c.cy = std::atof(v.c_str());
        else if (k == "--cz") c.cz = std::atof(v.c_str());
        else if (k == "--gridSize") { grid = std::atoi(v.c_str()); has_grid = true; }
        else if (k == "--timeSteps") { steps = std::atoi(v.c_str()); has_steps = true; }
        else if (k == "--endTime") { end = std::atof(v.c_str()); has_end = true; }
        else if (k == "--cfl") { cfl = std::atof(v.c_str()); has_cfl = true; }
        else if (k == "--order") order = std::atoi(v.c_str());
        else {
            std::cerr << "unrecognised option " << k << "\n";
            usage(1);
        }
    }
    if (order == 1) c.rk = false;
    else if (order == 4) c.rk = true;
    else {
        std::cerr << "Order " << order << " is not supported\nAborting.\n";
        std::exit(-1);
    }
    if (has_end) {
        c.endtime = end;
        c.dt = c.endtime / c.timesteps;
    }
    if (has_cfl && has_grid && has_steps) {
        std::cerr << "Error: setting cfl, gridSize and timeSteps not allowed\nAborting\n";
        std::exit(1);
    }
    if (has_grid) {
        c.gridsize = grid;
        c.dx = 1. / c.gridsize;
    }
    if (has_steps) {
        c.timesteps = steps;
        c.dt = c.endtime / c.timesteps;
    }
    if (has_cfl) {
        if (has_grid) {
            const double dt = c.dx * c.dx * cfl / (1.5 * c.nu0);
            c.timesteps = (int)(c.endtime / dt + .5);
            c.dt = c.endtime / c.timesteps;
        } else {
            // the reference derives the grid without the 3/2 factor of its CFL definition
            const double dx = std::sqrt(c.dt * c.nu0 / cfl);
            c.gridsize = (int)(1. / dx + .5);
            c.dx = 1. / c.gridsize;
        }
    }
    c.cfl = c.dt * 1.5 * c.nu0 / (c.dx * c.dx);
    return c;
}

double exactQ(const HeatConfiguration& c, double x, double y, double z, double t)
{
    return sp::exactQ(c.nu0, c.nufreq, c.cx, c.cy, c.cz, x, y, z, t);
}

void fillQ(ConvectionField& q, const HeatConfiguration& c, double t)
{
    sp::fillQ(q, c.nu0, c.nufreq, c.cx, c.cy, c.cz, t, 0., 1., 0., 1., 0., 1.);
}

// computeError, src/Convection/ConvectionSolution.h:55-103
double computeError(const ConvectionField& q, const HeatConfiguration& c, double t, double& exact_inf,
                    double& error_inf, ConvectionField* errfield)
{
    q.SynchronizeHostStorage();
    const IJKSize& d = q.calculationDomain();
    const double hx = 1. / d.iSize() / 2., hy = 1. / d.jSize() / 2., hz = 1. / d.kSize() / 2.;
    exact_inf = error_inf = 0.;
    for (int i = 0; i < d.iSize(); ++i)
        for (int j = 0; j < d.jSize(); ++j)
            for (int k = 0; k < d.kSize(); ++k) {
                const double ex = exactQ(c, (2 * i + 1) * hx, (2 * j + 1) * hy, (2 * k + 1) * hz, t);
                exact_inf = std::max(exact_inf, std::abs(ex));
                const double e = q(i, j, k) - ex;
                error_inf = std::max(error_inf, std::abs(e));
                if (errfield) (*errfield)(i, j, k) = e;
            }
    return error_inf / exact_inf;
}

}  // namespace

int main(int argc, char** argv)
{
    const HeatConfiguration conf = parse(argc, argv);
    const double h = 1. / conf.gridsize;
    std::cout << "convection: 3-D advection-diffusion on the periodic unit cube, one GPU\n"
              << "  grid " << conf.gridsize << "^3 (dx = " << conf.dx << "), " << conf.timesteps
              << " steps of dt = " << conf.dt << " to T = " << conf.endtime << ", CFL " << conf.cfl << "\n"
              << "  nu0 = " << conf.nu0 << ", nufreq = " << conf.nufreq << ", c = (" << conf.cx << ", " << conf.cy
              << ", " << conf.cz << ")\n"
              << "  solver: " << (conf.rk ? "Runge-Kutta 4" : "explicit Euler") << "\n"
              << "  domain [0, " << conf.gridsize * h << "]^3 on one process\n\n";

    try {
        IJKSize domain;
        domain.Init(conf.gridsize, conf.gridsize, conf.gridsize);
        KBoundary kb;
        kb.Init(-2, 2);
        ConvectionField qIn, qOut, errfield, exactfield;
        qIn.Init("q", domain, kb);
        qOut.Init("q", domain, kb);
        errfield.Init("error", domain, kb);
        exactfield.Init("exact", domain, kb);

        fillQ(qIn, conf, 0.);
        qIn.SynchronizeDeviceStorage();
        {
            MatFile initmat("init.mat");
            initmat.addField("qIn", qIn);
        }
        Convection convection(conf.gridsize, conf.gridsize, conf.gridsize, conf.dx, conf.nu0, conf.nufreq, conf.cx,
                              conf.cy, conf.cz);
        MatFile* mat = nullptr;
        if (conf.mat) {
            // the reference declares a cell of `timesteps` entries but stores only the
            // initial state; this cell holds exactly that one entry
            mat = new MatFile("convection_0_0_0.mat");
            mat->startCell("q", 1);
            mat->addField(qIn, 0);
        }

        check(sp_ctx_sync(context()));
        const auto t0 = std::chrono::steady_clock::now();
        if (conf.rk)
            convection.DoRK4(qIn, qOut, 0., conf.dt, conf.timesteps);
        else
            convection.DoEuler(qIn, qOut, 0., conf.dt, conf.timesteps);
        check(sp_ctx_sync(context()));
        const double e = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

        if (mat) {
            mat->endCell();
            delete mat;
        }

        double exact_inf = 0., error_inf = 0.;
        computeError(qOut, conf, conf.endtime, exact_inf, error_inf, conf.mat ? &errfield : nullptr);
        std::cout << "Timing results:\n"
                  << " - total time: " << e * 1000. << " msec\n"
                  << "\n"
                  << "Errors on all processes:\n"
                  << " - Process 0 has exact=" << exact_inf << ", error=" << error_inf << "\n"
                  << "\n Total relative error: " << error_inf / exact_inf << "\n";

        if (conf.mat) {
            fillQ(exactfield, conf, conf.endtime);
            // both were filled on the host; addField reads the device copy
            exactfield.SynchronizeDeviceStorage();
            errfield.SynchronizeDeviceStorage();
            MatFile matfile("result_0_0_0.mat");
            matfile.addField(qIn, -1);
            matfile.addField(qOut, -1);
            matfile.addField("error", errfield, -1);
            matfile.addField("exact", exactfield, -1);
        }
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
