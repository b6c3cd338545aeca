// TotalLagrangianHIPT.h -- element group of the continuum-solid hot path on
// one MI355X.
//
// Mirrors the element-group side of the reference's plugin interface
// (ElementBaseT, tahoe/src/elements/ElementBaseT.h:65-480) for the two
// formulations of the path:
//   total_lagrangian        -> TotalLagrangianT (TotalLagrangianT.cpp)
//   updated_lagrangian_Q1P0 -> SimoQ1P0 (SimoQ1P0.cpp)
// with Simo_isotropic materials (SimoIso3D.cpp) on 8-node hexahedra with
// 2x2x2 Gauss points.  The element group owns the device problem (a
// HIPContext); the global matrix and the nonlinear solver share it.
//
// Method names and meaning follow ElementBaseT: FormRHS / FormLHS
// (ElementBaseT.cpp:114-160), Equations (:196-219), InitStep / CloseStep /
// ResetStep (ElementBaseT.h:176-185), TangentType (:93).  Errors throw
// ExceptionT with the reference's codes (kBadJacobianDet for J <= 0).
#ifndef TAHOE_HIP_TOTAL_LAGRANGIAN_HIP_T_H
#define TAHOE_HIP_TOTAL_LAGRANGIAN_HIP_T_H

#include <map>
#include <vector>

#include "ExceptionT.h"
#include "HIPContext.h"

namespace tahoe_hip {

class TotalLagrangianHIPT {
public:
    enum FormulationT { kTotalLagrangian = TG_TOTAL_LAGRANGIAN, kQ1P0 = TG_Q1P0 };
    // GlobalT::SystemTypeT values the element group can report
    enum SystemTypeT { kUndefined = -1, kDiagonal = 0, kSymmetric = 1, kNonSymmetric = 2 };

    TotalLagrangianHIPT(HIPContext& ctx, FormulationT formulation);

    // ---- parameters (TakeParameterList) ------------------------------------
    // nodal coordinates [nnd][3], 0-based connectivity [nel][8] and the
    // element block (0-based material index) of every element
    void SetGeometry(const std::vector<double>& X, const std::vector<int>& conn, const std::vector<int>& block);
    // Simo_isotropic with E_and_nu (IsotropicT::Set_E_nu, IsotropicT.cpp:32-45)
    void SetMaterial(int block, double E, double nu);
    // kinematic BC cards: node, dof (0-based) and value at schedule 1
    void SetKinematicBC(const std::vector<int>& nodes, const std::vector<int>& dofs,
                        const std::vector<double>& values);
    void SetPrecision(int precision) { fCtx.Check(tg_set_precision(fCtx.Handle(), precision), "SetPrecision"); }
    // material model of every block: TG_MAT_SIMO_ISOTROPIC or TG_MAT_STVENANT
    void SetMaterialModel(int model)
    {
        fCtx.Check(tg_set_material_model(fCtx.Handle(), model), "SetMaterialModel");
    }
    // nodal force cards (FieldT's FBC_CardT list); ApplyForces sets their
    // current values (value x schedule value, FBC_CardT::CurrentValue)
    void SetForces(const std::vector<int>& nodes, const std::vector<int>& dofs, const std::vector<double>& values);
    void ApplyForces(const std::vector<double>& current_values);

    // equation numbering and device storage (FEManagerT::SetEquationSystem)
    int InitializeEquations();

    // ---- ElementBaseT -----------------------------------------------------
    SystemTypeT TangentType() const { return kSymmetric; }
    // FieldT::Equations: [nnd][3], 1-based, -1 prescribed
    const std::vector<int>& Equations() const { return fEqnos; }
    int NumEquations() const { return fNumEq; }
    int NumNodes() const { return fNumNodes; }
    int NumElements() const { return fNumElements; }
    FormulationT Formulation() const { return fFormulation; }

    // residual R = -f_int at the current displacement; R [n_eq] may be null
    void FormRHS(std::vector<double>* R, double* norm);
    // stiffness into the device MSR matrix the HIPMSRMatrixT reads
    void FormLHS();

    // prescribed values at time t (nStaticIntegrator::ConsistentKBC)
    void InitStep(double time);
    void CloseStep();
    void ResetStep();

    // nodal displacements [nnd][3] (FieldT::Update / LastUpdate)
    void SetDisplacement(const std::vector<double>& u);
    std::vector<double> Displacement() const;

    HIPContext& Context() { return fCtx; }

private:
    HIPContext& fCtx;
    FormulationT fFormulation;
    int fNumNodes = 0, fNumElements = 0, fNumEq = 0;
    std::map<int, std::pair<double, double>> fMaterials;  // block -> (E, nu)
    std::vector<int> fEqnos;
    std::vector<int> fForceNodes, fForceDofs;
};

}  // namespace tahoe_hip

#endif
