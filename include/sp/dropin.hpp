// dropin.hpp -- the whole C++ host API under the reference's unqualified names.
//
// Code written against the reference (src/Parareal/main.cpp, src/Convection/main.cpp)
// names its types without a namespace: RuntimeConfiguration, ConvectionField, Convection,
// Parareal<...>, IJKSize, KBoundary, MatFile, fillQ, ...  Including this header in place of
// the reference's RuntimeConfiguration.h, ConvectionSolution.h, Convection.h,
// SharedInfrastructure.h, StencilFramework.h, MatFile.h and Parareal.h lets such code
// compile with no other change.  MPI stays the caller's own: the Parareal constructor takes
// any communicator value (sp::Comm).
#pragma once

#include "Convection.hpp"
#include "ConvectionField.hpp"
#include "ConvectionSolution.hpp"
#include "MatFile.hpp"
#include "Parareal.hpp"
#include "RuntimeConfiguration.hpp"

using sp::Convection;
using sp::ConvectionField;
using sp::convectionBoundaryLines;
using sp::exactQ;
using sp::fillQ;
using sp::IJKBoundary;
using sp::IJKSize;
using sp::KBoundary;
using sp::MatFile;
using sp::ModeCompare;
using sp::ModeParallel;
using sp::ModeSerial;
using sp::ModeTiming;
using sp::Parareal;
using sp::RunMode;
using sp::RuntimeConfiguration;
