// Compile check of the drop-in C++ API: a driver written the way src/Parareal/main.cpp
// (lines 142-300) calls the reference -- unqualified type names, the six-argument Parareal
// constructor with the rank's start time and MPI_COMM_WORLD, the field's storage and
// boundary accessors through a computeErrorReference of the reference's shape, fillQ with
// the box bounds, MatFile output, SynchronizeHost / SynchronizeDevice helpers -- with the
// reference's includes replaced by sp/dropin.hpp and nothing else changed.  Built by
// tests/test_dropin_compile.py with g++ -c (no link, no GPU); MPI comes from the image's
// mpi.h.
#include <cmath>
#include <cstdio>
#include <iostream>

#include "mpi.h"
#include "sp/dropin.hpp"

template <typename TDataField>
inline void SynchronizeHost(TDataField& field)
{
#ifdef __CUDA_BACKEND__
    field.SynchronizeHostStorage();
#endif
}

template <typename TDataField>
inline void SynchronizeDevice(TDataField& field)
{
#ifdef __CUDA_BACKEND__
    field.SynchronizeDeviceStorage();
#endif
}

inline void SynchronizeCUDA() {}

// max |field - reference| / max |field| over the interior, walking the field's boundary
// offsets as the reference's helper does
template <typename TDataField>
double computeErrorReference(const TDataField& field, const TDataField& reference)
{
    const IJKSize& size = field.calculationDomain();
    const IJKBoundary boundary = field.boundary();
    const int istart = -boundary.iMinusOffset();
    (void)istart;
    double error = 0., maxfield = 0.;
    for (int i = 0; i < size.iSize(); ++i)
        for (int j = 0; j < size.jSize(); ++j)
            for (int k = 0; k < size.kSize(); ++k) {
                error = std::max(error, std::abs(field(i, j, k) - reference(i, j, k)));
                maxfield = std::max(std::abs(field(i, j, k)), maxfield);
            }
    return error / maxfield;
}

int main(int argc, char** argv)
{
    MPI_Init(&argc, &argv);
    int commsize, commrank;
    MPI_Comm_size(MPI_COMM_WORLD, &commsize);
    MPI_Comm_rank(MPI_COMM_WORLD, &commrank);
    const bool isLast = commrank == commsize - 1;

    RuntimeConfiguration conf(argc, argv);
    const double timeStart = conf.timeSliceSize() * commrank;

    IJKSize domain;
    domain.Init(conf.gridSize(), conf.gridSize(), conf.gridSize());
    KBoundary kboundary;
    kboundary.Init(-convectionBoundaryLines, convectionBoundaryLines);

    ConvectionField q, qinitial;
    q.Init("q", domain, kboundary);
    qinitial.Init("qinitial", domain, kboundary);
    Convection convection(conf.gridSize(), conf.gridSize(), conf.gridSize(), conf.dx(), conf.nu0(), conf.nufreq(),
                          conf.cx(), conf.cy(), conf.cz());

    Parareal<Convection, ConvectionField> parareal(convection, qinitial, q, timeStart, conf, MPI_COMM_WORLD);

    // what Parareal.h reads from its fields
    const IJKBoundary& b = qinitial.boundary();
    KBoundary kb;
    kb.Init(b.kMinusOffset(), b.kPlusOffset());
    double* pRecv = qinitial.storage().pStorageBase();
    const IJKSize& psize = q.storage().paddedSize();
    const int dataSize = psize.iSize() * psize.jSize() * psize.kSize();
    (void)pRecv;
    (void)dataSize;

    if (conf.mode() == ModeCompare) {
        const int tauSamples = 4;
        double tauF = MPI_Wtime();
        convection.DoRK4(qinitial, qinitial, 0., conf.dtFine(), tauSamples * conf.timeStepsFinePerTimeSlice());
        SynchronizeCUDA();
        tauF = MPI_Wtime() - tauF;
        double tauG = MPI_Wtime();
        convection.DoEuler(qinitial, qinitial, 0., conf.dtCoarse(), tauSamples * conf.timeStepsCoarsePerTimeSlice());
        SynchronizeCUDA();
        tauG = MPI_Wtime() - tauG;
        const double tauRatio = tauG / tauF;

        SynchronizeHost(qinitial);
        fillQ(qinitial, conf.nu0(), conf.nufreq(), conf.cx(), conf.cy(), conf.cz(), 0., 0., 1., 0., 1., 0., 1.);
        SynchronizeDevice(qinitial);

        MPI_Barrier(MPI_COMM_WORLD);
        double eserial = MPI_Wtime();
        parareal.DoSerial();
        eserial = MPI_Wtime() - eserial;

        ConvectionField qreference = q;
        SynchronizeHost(qreference);

        SynchronizeHost(qinitial);
        fillQ(qinitial, conf.nu0(), conf.nufreq(), conf.cx(), conf.cy(), conf.cz(), 0., 0., 1., 0., 1., 0., 1.);
        SynchronizeDevice(qinitial);

        MPI_Barrier(MPI_COMM_WORLD);
        double eparallel = MPI_Wtime();
        parareal.DoParallel();
        eparallel = MPI_Wtime() - eparallel;

        MPI_Barrier(MPI_COMM_WORLD);
        if (isLast) {
            const double e = computeErrorReference(q, qreference);
            std::cout << "Serial run time: " << eserial << "\nParallel run time: " << eparallel
                      << "\nSpeedup: " << eserial / eparallel << "\nError at end: " << e << "\n"
                      << "tau ratio: " << tauRatio << std::endl;
            MatFile matfile("result.mat");
            matfile.addField("q", q);
            matfile.addField("qreference", qreference);
        }
    } else if (conf.mode() == ModeTiming) {
        SynchronizeHost(qinitial);
        fillQ(qinitial, conf.nu0(), conf.nufreq(), conf.cx(), conf.cy(), conf.cz(), 0., 0., 1., 0., 1., 0., 1.);
        SynchronizeDevice(qinitial);
        std::vector<double> times;
        parareal.DoTimedParallel(times);
        for (double t : times) std::printf("%e\n", t);
    }

    MPI_Finalize();
    return 0;
}
