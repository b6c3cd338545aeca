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
        if (fPar.method == kNonlinearCG) return SolveCG(max_iterations);
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

// ---- PCG_solver (PCGSolver_LS) ---------------------------------------------

static double dot(const std::vector<double>& a, const std::vector<double>& b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

// DiagonalMatrixT::Factorize/BackSubstitute (DiagonalMatrixT.cpp:267-320):
// v_i *= 1/d_i, or *= d_i itself where |d_i| <= kSmall = 1e-12
void NLSolverHIP::Precondition(std::vector<double>& v) const
{
    for (size_t i = 0; i < v.size(); i++) v[i] *= fMinv[i];
}

// NLSolver::Solve (NLSolver.cpp:57-264) with PCGSolver_LS::Iterate
// (PCGSolver_LS.cpp:108-119): the diagonal is re-formed on the first
// iteration and every `restart` iterations after it
NLSolverHIP::SolutionStatusT NLSolverHIP::SolveCG(int max_iterations)
{
    const int reform = std::max(fPar.restart, 1);
    std::vector<double> R;
    double error = 0.0;
    fRestart_count = -1;
    const int restart_iteration = fNumIteration;  // NLSolver::InitStep's marker
    fElements.FormRHS(&R, &error);
    SolutionStatusT flag = ExitIteration(error, fNumIteration);
    int num_iterations = 0, tan_iterations = 0;
    while (flag == kContinue && (max_iterations == -1 || num_iterations < max_iterations)) {
        num_iterations++;
        tan_iterations++;
        if (num_iterations == 1 || tan_iterations >= reform) {
            tan_iterations = 0;
            fElements.FormLHS();
            std::vector<double> d;
            fLHS.CopyDiagonal(d);
            fMinv.resize(d.size());
            for (size_t i = 0; i < d.size(); i++) fMinv[i] = std::fabs(d[i]) > 1.0e-12 ? 1.0 / d[i] : d[i];
        }
        // CGSearch: restart on the step's first iteration, at restart_count
        // == restart, or at the first iteration after a relaxation marker
        const bool start_relaxation = restart_iteration == fNumIteration;
        std::vector<double> R0 = R;
        fRestart_count++;
        if (fRestart_count == 0 || fRestart_count == fPar.restart || start_relaxation) {
            fR_last = R;
            Precondition(R);
            fu_last = R;
            fRestart_count = 0;
        } else
            CGSearch(R);
        LineSearch(R, R0);
        fNumIteration++;
        fElements.FormRHS(&R, &error);
        fLastError = error;
        flag = ExitIteration(error, fNumIteration);
    }
    if (fNumIteration == -1) fLastError = error;
    return flag;
}

// PCGSolver_LS::CGSearch, scaled Polak-Ribiere branch (Bertsekas 6.32):
// beta = R.M^-1(R - R_last) / R_last.M^-1 R_last, steepest descent when the
// denominator vanishes; d = M^-1 R + beta d_last
void NLSolverHIP::CGSearch(std::vector<double>& R)
{
    std::vector<double> diff(R.size());
    for (size_t i = 0; i < R.size(); i++) diff[i] = R[i] - fR_last[i];
    Precondition(diff);
    double beta = dot(R, diff);
    diff = fR_last;
    Precondition(diff);
    double den = dot(fR_last, diff);
    if (std::fabs(den) < 1.0e-24) {
        den = 1.0;
        beta = 0.0;
    } else
        beta /= den;
    fR_last = R;
    Precondition(R);
    for (size_t i = 0; i < R.size(); i++) R[i] += beta * fu_last[i];
    fu_last = R;
}

// G(s) = d . R(u + s d), moving the field from the last evaluated step
// (PCGSolver_LS::GValue, PCGSolver_LS.cpp:396-410)
double NLSolverHIP::GValue(double step)
{
    const std::vector<int>& eq = fElements.Equations();
    std::vector<double> u = fElements.Displacement();
    const double ds = step - fStepCurrent;
    for (size_t k = 0; k < eq.size(); k++)
        if (eq[k] > 0) u[k] += ds * fUpdate[(size_t)eq[k] - 1];
    fElements.SetDisplacement(u);
    fStepCurrent = step;
    double err = 0.0;
    fElements.FormRHS(&fRwork, &err);
    return dot(fUpdate, fRwork);
}

// PCGSolver_LS::Update (PCGSolver_LS.cpp:216-393): secant iterations on G
// between s_a = 0 (G = d . R0) and s_b = 1, bounded by [0, max_step], until
// |G| <= abs_tolerance, |G/G_0| <= line_search_tolerance or the iteration
// budget; on giving up, the evaluated step of smallest |G| (the reference's
// selection rule) is taken
void NLSolverHIP::LineSearch(const std::vector<double>& dir, const std::vector<double>& R0)
{
    fUpdate = dir;
    fStepCurrent = 0.0;
    if (fPar.search_iterations == 0) {
        GValue(1.0);
        return;
    }
    const int nmax = std::max(fPar.search_iterations, 3);
    std::vector<double> ds((size_t)nmax + 1, 0.0), dg((size_t)nmax + 1, 0.0);
    double s_a = 0.0, G_a = dot(fUpdate, R0);
    ds[0] = s_a;
    dg[0] = G_a;
    double s_b = 1.0, G_b = GValue(s_b);
    ds[1] = s_b;
    dg[1] = G_b;
    const double G_0 = std::fabs(G_a) > std::fabs(G_b) ? G_b : G_a;
    int count = 2;
    double G_new = G_b;
    bool give_up = false;
    do {
        const double m = (G_a - G_b) / (s_a - s_b);
        const double b = G_b - m * s_b;
        double s_new = -b / m;
        if (s_new > fPar.max_step || s_new < 0.0 || !std::isfinite(s_new)) {
            give_up = true;
            if (s_new > fPar.max_step) {
                s_new = fPar.max_step;
                G_new = GValue(s_new);
                ds[count] = s_new;
                dg[count] = G_new;
                count++;
            }
            break;
        }
        G_new = GValue(s_new);
        ds[count] = s_new;
        dg[count] = G_new;
        if (std::fabs(G_a) > std::fabs(G_new) && std::fabs(G_a) > std::fabs(G_b)) {
            G_a = G_new;
            s_a = s_new;
            give_up = false;
        } else if (std::fabs(G_b) > std::fabs(G_new) && std::fabs(G_b) > std::fabs(G_a)) {
            G_b = G_new;
            s_b = s_new;
            give_up = false;
        } else if (G_b * G_a > 0) {
            if (G_a * G_new < 0) {
                G_a = G_new;
                s_a = s_new;
            } else if (G_b * G_new < 0) {
                G_b = G_new;
                s_b = s_new;
            } else
                give_up = true;
        } else
            give_up = true;
        if (++count >= fPar.search_iterations) give_up = true;
    } while (std::fabs(G_new) > fPar.abs_tolerance && std::fabs(G_new / G_0) > fPar.orthog_tolerance && !give_up);
    if (give_up) {
        const double kSmall = 1.0e-12;
        double s_best = std::fabs(ds[0]), G_best = std::fabs(dg[0]);
        int best = 0;
        for (int i = 1; i < count; i++) {
            const double s_test = std::fabs(ds[i]), G_test = std::fabs(dg[i]);
            if (std::fabs(s_best) < kSmall || (s_test > kSmall && G_test < G_best)) {
                s_best = s_test;
                G_best = G_test;
                best = i;
            }
        }
        GValue(ds[best]);
    }
}

}  // namespace tahoe_hip
