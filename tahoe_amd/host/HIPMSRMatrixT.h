// HIPMSRMatrixT.h -- the global matrix of the hot path, resident on the GPU.
//
// Mirrors GlobalMatrixT (tahoe/src/primitives/globalmatrix/GlobalMatrixT.h:24-223)
// as MSRMatrixT implements it (MSRMatrixT.cpp): upper-triangle MSR storage in
// the reference's global equation numbering (EquationNumberScope kGlobal, no
// renumbering), Multx and CopyDiagonal on the stored values, and Solve =
// Factorize + BackSubstitute with the factorisation replaced by Jacobi-PCG.
// The values are assembled on the device by the element group
// (TotalLagrangianHIPT::FormLHS), so the element-matrix Assemble overloads
// have nothing to do.  Host copies of bindx/val exist only on request
// (parity checks, PrintLHS).
#ifndef TAHOE_HIP_HIP_MSR_MATRIX_T_H
#define TAHOE_HIP_HIP_MSR_MATRIX_T_H

#include <ostream>
#include <vector>

#include "HIPContext.h"

namespace tahoe_hip {

class HIPMSRMatrixT {
public:
    enum CheckCodeT { kNoCheck = 0, kZeroPivots = 1, kAllPivots = 2, kPrintLHS = 3, kPrintRHS = 4,
                      kPrintSolution = 5, kCheckLHS = 6 };
    enum EquationNumberScopeT { kLocal = 0, kGlobal = 1 };

    // rtol / maxit: the Jacobi-PCG that stands in for the factorisation
    HIPMSRMatrixT(HIPContext& ctx, double rtol = 1e-10, int maxit = 100000, int check_code = kNoCheck);

    // GlobalMatrixT::Initialize(tot_num_eq, loc_num_eq, start_eq): one
    // process, so tot = loc and start = 1 (GlobalMatrixT.cpp:41-71)
    void Initialize(int tot_num_eq, int loc_num_eq, int start_eq);
    void Clear() {}
    EquationNumberScopeT EquationNumberScope() const { return kGlobal; }
    bool RenumberEquations() const { return false; }
    bool SolvePreservesData() const { return true; }
    int CheckCode() const { return fCheckCode; }

    // Solve(rhs): rhs holds b on entry and x on exit; false when the PCG
    // does not reach its tolerance (GlobalMatrixT::Solve returns false on
    // a caught exception, GlobalMatrixT.cpp:77-114)
    bool Solve(std::vector<double>& rhs);
    // b = K x (MSRMatrixT::Multx, MSRMatrixT.cpp:385-420)
    bool Multx(const std::vector<double>& x, std::vector<double>& b) const;
    // diagonal of K (MSRMatrixT::CopyDiagonal, MSRMatrixT.cpp:351-378)
    bool CopyDiagonal(std::vector<double>& diags) const;

    // MSR arrays (MSRBuilderT::GenerateMSR; MSRMatrixT::fval)
    std::vector<int> BIndex() const;
    std::vector<double> Values() const;
    long MSRLength() const { return fMSRLength; }
    int NumEquations() const { return fNumEq; }

    // the upper triangle as 1-based RCV text, exact zeros dropped, 15
    // significant digits (MSRMatrixT::PrintLHS, MSRMatrixT.cpp:449-471)
    void PrintLHS(std::ostream& out) const;

    int LastIterations() const { return fLastIterations; }
    double LastResidual() const { return fLastResidual; }

private:
    HIPContext& fCtx;
    double fRTol;
    int fMaxIt;
    int fCheckCode;
    int fNumEq = 0;
    long fMSRLength = 0;
    int fLastIterations = 0;
    double fLastResidual = 0.0;
};

}  // namespace tahoe_hip

#endif
