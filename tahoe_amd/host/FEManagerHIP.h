// FEManagerHIP.h -- static time loop of one input deck on the GPU.
//
// Mirrors FEManagerT::Solve's static path (tahoe/src/main/FEManagerT.cpp:138-269)
// for one solution group: for step n = 1..num_steps at t = n * time_step,
// InitStep (prescribed values), NLSolver::Solve, then CloseStep on
// convergence; a failed step stops the run with kGeneralFail.  Nodal
// displacements of every step go to a text file laid out like the
// reference's ".io0.run" output (OutputBaseT, nodal values D_X D_Y D_Z).
#ifndef TAHOE_HIP_FE_MANAGER_HIP_H
#define TAHOE_HIP_FE_MANAGER_HIP_H

#include <ostream>
#include <string>
#include <vector>

#include "InputDeck.h"
#include "NLSolverHIP.h"

namespace tahoe_hip {

struct StepRecord {
    int step;
    double time;
    int iterations;       // Newton corrections
    long linear_iterations;
    double error0, error;
    std::vector<double> u;  // [nnd][3]
};

class FEManagerHIP {
public:
    struct Options {
        int device = 0;
        int precision = TG_FP64;
        NLSolverHIP::DriverT driver = NLSolverHIP::kDevice;
        double linear_tolerance = -1.0;  // < 0: from the deck
    };

    FEManagerHIP(const InputDeck& deck, const Options& opt);

    // runs every step; throws ExceptionT(kGeneralFail) when a step fails
    const std::vector<StepRecord>& Solve(std::ostream* log = nullptr);
    // digits after the point of the %e fields (the reference prints 12)
    void WriteRun(std::ostream& out, int digits = 12) const;

    TotalLagrangianHIPT& Elements() { return fElements; }
    HIPMSRMatrixT& Matrix() { return fLHS; }

private:
    const InputDeck& fDeck;
    Options fOpt;
    HIPContext fCtx;
    TotalLagrangianHIPT fElements;
    HIPMSRMatrixT fLHS;
    int fSchedule = -1;
    std::vector<StepRecord> fSteps;
};

}  // namespace tahoe_hip

#endif
