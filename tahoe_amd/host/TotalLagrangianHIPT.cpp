// TotalLagrangianHIPT.cpp -- element group forwarding to libtahoe_gpu.
#include "TotalLagrangianHIPT.h"

#include <cmath>

namespace tahoe_hip {

TotalLagrangianHIPT::TotalLagrangianHIPT(HIPContext& ctx, FormulationT formulation)
    : fCtx(ctx), fFormulation(formulation)
{
    fCtx.Check(tg_set_formulation(fCtx.Handle(), (int)formulation), "TotalLagrangianHIPT");
}

void TotalLagrangianHIPT::SetGeometry(const std::vector<double>& X, const std::vector<int>& conn,
                                      const std::vector<int>& block)
{
    const char* caller = "TotalLagrangianHIPT::SetGeometry";
    if (X.size() % 3 != 0 || conn.size() % 8 != 0) ExceptionT::SizeMismatch(caller, "expecting [nnd][3] and [nel][8]");
    fNumNodes = (int)(X.size() / 3);
    fNumElements = (int)(conn.size() / 8);
    if (!block.empty() && block.size() != (size_t)fNumElements)
        ExceptionT::SizeMismatch(caller, "one block index per element");
    fCtx.Check(tg_set_mesh(fCtx.Handle(), fNumNodes, fNumElements, X.data(), conn.data(),
                           block.empty() ? nullptr : block.data()),
               caller);
}

void TotalLagrangianHIPT::SetMaterial(int block, double E, double nu)
{
    // IsotropicT::Set_E_nu (IsotropicT.cpp:32-45): the admissible range of
    // Poisson's ratio for a 3D solid
    if (E <= 0.0 || nu <= -1.0 || nu >= 0.5)
        ExceptionT::BadInputValue("TotalLagrangianHIPT::SetMaterial", "E > 0 and -1 < nu < 1/2 required");
    const double mu = 0.5 * E / (1.0 + nu);
    const double kappa = E / (3.0 * (1.0 - 2.0 * nu));
    fMaterials[block] = {E, nu};
    fCtx.Check(tg_set_material(fCtx.Handle(), block, mu, kappa), "TotalLagrangianHIPT::SetMaterial");
}

void TotalLagrangianHIPT::SetKinematicBC(const std::vector<int>& nodes, const std::vector<int>& dofs,
                                         const std::vector<double>& values)
{
    if (nodes.size() != dofs.size() || nodes.size() != values.size())
        ExceptionT::SizeMismatch("TotalLagrangianHIPT::SetKinematicBC");
    fCtx.Check(tg_set_kbc(fCtx.Handle(), (int)nodes.size(), nodes.data(), dofs.data(), values.data()),
               "TotalLagrangianHIPT::SetKinematicBC");
}

void TotalLagrangianHIPT::SetForces(const std::vector<int>& nodes, const std::vector<int>& dofs,
                                    const std::vector<double>& values)
{
    if (nodes.size() != dofs.size() || nodes.size() != values.size())
        ExceptionT::SizeMismatch("TotalLagrangianHIPT::SetForces");
    fForceNodes = nodes;
    fForceDofs = dofs;
    fCtx.Check(tg_set_forces(fCtx.Handle(), (int)nodes.size(), nodes.data(), dofs.data(), values.data()),
               "TotalLagrangianHIPT::SetForces");
}

void TotalLagrangianHIPT::ApplyForces(const std::vector<double>& current_values)
{
    if (current_values.size() != fForceNodes.size()) ExceptionT::SizeMismatch("TotalLagrangianHIPT::ApplyForces");
    if (fForceNodes.empty()) return;
    fCtx.Check(tg_set_forces(fCtx.Handle(), (int)fForceNodes.size(), fForceNodes.data(), fForceDofs.data(),
                             current_values.data()),
               "TotalLagrangianHIPT::ApplyForces");
    fCtx.Check(tg_apply_forces(fCtx.Handle(), 1.0), "TotalLagrangianHIPT::ApplyForces");
}

int TotalLagrangianHIPT::InitializeEquations()
{
    long len = 0;
    fCtx.Check(tg_initialize(fCtx.Handle(), &fNumEq, &len), "TotalLagrangianHIPT::InitializeEquations");
    fEqnos.assign((size_t)fNumNodes * 3, 0);
    fCtx.Check(tg_get_eqnos(fCtx.Handle(), fEqnos.data()), "TotalLagrangianHIPT::Equations");
    return fNumEq;
}

void TotalLagrangianHIPT::FormRHS(std::vector<double>* R, double* norm)
{
    if (R) R->assign((size_t)fNumEq, 0.0);
    fCtx.Check(tg_form_rhs(fCtx.Handle(), R ? R->data() : nullptr, norm), "TotalLagrangianHIPT::FormRHS");
}

void TotalLagrangianHIPT::FormLHS() { fCtx.Check(tg_form_lhs(fCtx.Handle()), "TotalLagrangianHIPT::FormLHS"); }

void TotalLagrangianHIPT::InitStep(double time)
{
    fCtx.Check(tg_apply_kbc(fCtx.Handle(), time), "TotalLagrangianHIPT::InitStep");
}

void TotalLagrangianHIPT::CloseStep() { fCtx.Check(tg_close_step(fCtx.Handle()), "TotalLagrangianHIPT::CloseStep"); }

void TotalLagrangianHIPT::ResetStep() { fCtx.Check(tg_reset_step(fCtx.Handle()), "TotalLagrangianHIPT::ResetStep"); }

void TotalLagrangianHIPT::SetDisplacement(const std::vector<double>& u)
{
    if (u.size() != (size_t)fNumNodes * 3) ExceptionT::SizeMismatch("TotalLagrangianHIPT::SetDisplacement");
    fCtx.Check(tg_set_displacement(fCtx.Handle(), u.data()), "TotalLagrangianHIPT::SetDisplacement");
}

std::vector<double> TotalLagrangianHIPT::Displacement() const
{
    std::vector<double> u((size_t)fNumNodes * 3);
    fCtx.Check(tg_get_displacement(fCtx.Handle(), u.data()), "TotalLagrangianHIPT::Displacement");
    return u;
}

}  // namespace tahoe_hip
