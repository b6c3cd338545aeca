// HIPMSRMatrixT.cpp -- GlobalMatrixT-shaped view of the device MSR matrix.
#include "HIPMSRMatrixT.h"

#include <cmath>
#include <iomanip>

namespace tahoe_hip {

HIPMSRMatrixT::HIPMSRMatrixT(HIPContext& ctx, double rtol, int maxit, int check_code)
    : fCtx(ctx), fRTol(rtol), fMaxIt(maxit), fCheckCode(check_code)
{
    // maxit 0: the linear solver's own limit (3 n_eq Arnoldi steps for GMRES,
    // NLSolver_NK's max_linear_iter = 0)
    if (!(rtol > 0.0) || maxit < 0) ExceptionT::BadInputValue("HIPMSRMatrixT", "linear tolerance and iteration limit");
}

void HIPMSRMatrixT::Initialize(int tot_num_eq, int loc_num_eq, int start_eq)
{
    const char* caller = "HIPMSRMatrixT::Initialize";
    if (tot_num_eq != loc_num_eq || start_eq != 1)
        ExceptionT::GeneralFail(caller, "one process owns every equation");
    int n = 0;
    fCtx.Check(tg_initialize(fCtx.Handle(), &n, &fMSRLength), caller);
    if (n != loc_num_eq) ExceptionT::SizeMismatch(caller, "equation count differs from the element group's");
    fNumEq = n;
}

bool HIPMSRMatrixT::Solve(std::vector<double>& rhs)
{
    if (rhs.size() != (size_t)fNumEq) ExceptionT::SizeMismatch("HIPMSRMatrixT::Solve");
    tg_lin_stats st{};
    const int rc = tg_solve(fCtx.Handle(), rhs.data(), fRTol, fMaxIt, &st);
    fLastIterations = st.iterations;
    fLastResidual = st.rel_residual;
    if (rc == TG_BAD_JACOBIAN) fCtx.Check(rc, "HIPMSRMatrixT::Solve");
    return rc == TG_OK && st.converged != 0;
}

bool HIPMSRMatrixT::Multx(const std::vector<double>& x, std::vector<double>& b) const
{
    if (x.size() != (size_t)fNumEq) return false;
    b.assign((size_t)fNumEq, 0.0);
    return tg_multx(fCtx.Handle(), x.data(), b.data()) == TG_OK;
}

bool HIPMSRMatrixT::CopyDiagonal(std::vector<double>& diags) const
{
    diags.assign((size_t)fNumEq, 0.0);
    return tg_copy_diagonal(fCtx.Handle(), diags.data()) == TG_OK;
}

std::vector<int> HIPMSRMatrixT::BIndex() const
{
    std::vector<int> b((size_t)fMSRLength);
    fCtx.Check(tg_get_msr_bindx(fCtx.Handle(), b.data()), "HIPMSRMatrixT::BIndex");
    return b;
}

std::vector<double> HIPMSRMatrixT::Values() const
{
    std::vector<double> v((size_t)fMSRLength);
    fCtx.Check(tg_get_msr_values(fCtx.Handle(), v.data()), "HIPMSRMatrixT::Values");
    return v;
}

void HIPMSRMatrixT::PrintLHS(std::ostream& out) const
{
    const std::vector<int> bindx = BIndex();
    const std::vector<double> val = Values();
    out << std::setprecision(14);
    for (int row = 0; row < fNumEq; row++) {
        if (std::fabs(val[row]) > 0.0) out << row + 1 << " " << row + 1 << " " << val[row] << '\n';
        for (int k = bindx[row]; k < bindx[row + 1]; k++)
            if (std::fabs(val[k]) > 0.0) out << row + 1 << " " << bindx[k] + 1 << " " << val[k] << '\n';
    }
}

}  // namespace tahoe_hip
