// NLSolverHIP.h -- Newton solver of one solution group on the GPU.
//
// Mirrors NLSolver (tahoe/src/solvers/NLSolver.cpp): Solve(max_iterations)
// returns kContinue/kConverged/kFailed after ExitIteration's tests
// (NLSolver.cpp:675-757): the first residual norm is the reference error;
// converged when ||R||/||R0|| < rel_tolerance or ||R|| < abs_tolerance,
// failed when ||R||/||R0|| > divergence_tolerance or the iteration count
// reaches max_iterations.  A kBadJacobianDet thrown by the element group
// makes Solve return kFailed, as NLSolver::Solve's catch does
// (NLSolver.cpp:250-263).
//
// Two drivers with the same tests:
//   kDevice -- the whole loop runs on the device (tg_newton); vectors never
//              leave HBM.
//   kHost   -- the reference's loop literally: FormRHS to the host, the
//              global matrix's Solve on the host vector, u += du, upload;
//              the plugins are used exactly as NLSolver uses them.
//
// Method kNonlinearCG is the reference's PCG_solver (PCGSolver_LS,
// PCGSolver_LS.cpp): NLSolver::Solve's loop with the tangent (here only its
// diagonal, DiagonalMatrixT) re-formed every `restart` iterations, search
// directions by preconditioned Polak-Ribiere CG restarted every `restart`
// iterations (CGSearch, :146-214), and a secant line search on
// G(s) = d . R(u + s d) (Update/GValue, :216-410).  It runs on the host
// driver: residuals come from the device, vectors are host arrays.
#ifndef TAHOE_HIP_NL_SOLVER_HIP_H
#define TAHOE_HIP_NL_SOLVER_HIP_H

#include "HIPMSRMatrixT.h"
#include "TotalLagrangianHIPT.h"

namespace tahoe_hip {

class NLSolverHIP {
public:
    enum SolutionStatusT { kContinue = 0, kConverged = 1, kFailed = 2 };
    enum DriverT { kDevice = 0, kHost = 1 };
    enum MethodT { kNewton = 0, kNonlinearCG = 1 };

    struct Parameters {
        int max_iterations = 20;
        int min_iterations = 0;
        double abs_tolerance = 1e-14;
        double rel_tolerance = 1e-10;
        double divergence_tolerance = 10.0;
        double linear_tolerance = 1e-10;
        int max_linear_iterations = 100000;
        // PCG_solver (PCGSolver_LS defaults, PCGSolver_LS.cpp:17-32)
        MethodT method = kNewton;
        int restart = 0;                 // CG restart and diagonal re-form interval
        int search_iterations = 3;       // line_search_iterations
        double orthog_tolerance = 0.25;  // line_search_tolerance
        double max_step = 2.5;           // max_step
    };

    NLSolverHIP(TotalLagrangianHIPT& elements, HIPMSRMatrixT& lhs, const Parameters& p, DriverT driver = kDevice);

    SolutionStatusT Solve(int max_iterations = -1);

    int IterationNumber() const { return fNumIteration; }
    double Error0() const { return fError0; }
    double LastError() const { return fLastError; }
    long LinearIterations() const { return fLinearIterations; }

private:
    SolutionStatusT ExitIteration(double error, int iteration);
    SolutionStatusT SolveDevice(int max_iterations);
    SolutionStatusT SolveHost(int max_iterations);
    SolutionStatusT SolveCG(int max_iterations);
    void CGSearch(std::vector<double>& R);
    void LineSearch(const std::vector<double>& dir, const std::vector<double>& R0);
    double GValue(double step);
    void Precondition(std::vector<double>& v) const;

    TotalLagrangianHIPT& fElements;
    HIPMSRMatrixT& fLHS;
    Parameters fPar;
    DriverT fDriver;
    int fNumIteration = -1;
    double fError0 = 0.0, fLastError = 0.0;
    long fLinearIterations = 0;
    // nonlinear CG state
    std::vector<double> fMinv, fR_last, fu_last, fUpdate, fRwork;
    int fRestart_count = -1;
    double fStepCurrent = 0.0;
};

}  // namespace tahoe_hip

#endif
