// NLSolverHIP.cpp -- NLSolver::Solve over the device element group and matrix.
#include "NLSolverHIP.h"

#include <cmath>

namespace tahoe_hip {

NLSolverHIP::NLSolverHIP(TotalLagrangianHIPT& elements, HIPMSRMatrixT& lhs, const Parameters& p, DriverT driver)
    : fElements(elements), fLHS(lhs), fPar(p), fDriver(driver)
{
    if (p.max_iterations < 0 || !(p.rel_tolerance > 0.0) || !(p.divergence_tolerance > 0.0))
        ExceptionT::BadInputValue("NLSolverHIP", "tolerances and iteration limit");
}

NLSolverHIP::SolutionStatusT NLSolverHIP::ExitIteration(double error, int iteration)
{
    if (iteration == -1) {
        fError0 = error;
        return fError0 < fPar.abs_tolerance ? kConverged : kContinue;
    }
    const double rel = error / fError0;
    if (rel > fPar.divergence_tolerance) return kFailed;
    if (iteration < fPar.min_iterations - 1) return kContinue;
    if (rel < fPar.rel_tolerance || error < fPar.abs_tolerance) return kConverged;
    if (iteration >= fPar.max_iterations) return kFailed;
    return kContinue;
}

NLSolverHIP::SolutionStatusT NLSolverHIP::Solve(int max_iterations)
{
    fNumIteration = -1;
    fLinearIterations = 0;
    try {
        return fDriver == kDevice ? SolveDevice(max_iterations) : SolveHost(max_iterations);
    } catch (const ExceptionT& e) {
        if (e.Code() == ExceptionT::kBadJacobianDet) return kFailed;
        throw;
    }
}

NLSolverHIP::SolutionStatusT NLSolverHIP::SolveDevice(int max_iterations)
{
    tg_newton_opts o;
    o.rel_tolerance = fPar.rel_tolerance;
    o.abs_tolerance = fPar.abs_tolerance;
    o.divergence_tolerance = fPar.divergence_tolerance;
    o.max_iterations = max_iterations == -1 ? fPar.max_iterations : std::min(max_iterations, fPar.max_iterations);
    o.lin_rtol = fPar.linear_tolerance;
    o.lin_maxit = fPar.max_linear_iterations;
    if (fPar.min_iterations > 0)
        ExceptionT::BadInputValue("NLSolverHIP::Solve", "min_iterations needs the host driver");
    tg_newton_stats st{};
    const int rc = tg_newton(fElements.Context().Handle(), &o, &st);
    fError0 = st.error0;
    fLastError = st.error;
    fNumIteration = st.iterations - 1;
    fLinearIterations = st.lin_iterations;
    if (rc == TG_BAD_JACOBIAN || st.status == TG_BAD_JACOBIAN) return kFailed;
    fElements.Context().Check(rc, "NLSolverHIP::Solve");
    return st.status == 0 ? kConverged : kFailed;
}

NLSolverHIP::SolutionStatusT NLSolverHIP::SolveHost(int max_iterations)
{
    const std::vector<int>& eq = fElements.Equations();
    std::vector<double> R;
    double error = 0.0;
    fElements.FormRHS(&R, &error);
    SolutionStatusT flag = ExitIteration(error, fNumIteration);
    int num_iterations = 0;
    while (flag == kContinue && (max_iterations == -1 || num_iterations < max_iterations)) {
        num_iterations++;
        // NLSolver::Iterate: form K, solve K du = R, update the field
        fElements.FormLHS();
        if (!fLHS.Solve(R)) ExceptionT::GeneralFail("NLSolverHIP::Solve", "linear solve failed");
        fLinearIterations += fLHS.LastIterations();
        std::vector<double> u = fElements.Displacement();
        for (size_t k = 0; k < eq.size(); k++)
            if (eq[k] > 0) u[k] += R[(size_t)eq[k] - 1];
        fElements.SetDisplacement(u);
        fNumIteration++;
        fElements.FormRHS(&R, &error);
        fLastError = error;
        flag = ExitIteration(error, fNumIteration);
    }
    if (fNumIteration == -1) fLastError = error;
    return flag;
}

}  // namespace tahoe_hip
