// FEManagerHIP.cpp -- deck -> element group, matrix and Newton loop.
#include "FEManagerHIP.h"

#include <cstdio>
#include <iomanip>

namespace tahoe_hip {

FEManagerHIP::FEManagerHIP(const InputDeck& deck, const Options& opt)
    : fDeck(deck),
      fOpt(opt),
      fCtx(opt.device),
      fElements(fCtx, deck.formulation == TG_Q1P0 ? TotalLagrangianHIPT::kQ1P0 : TotalLagrangianHIPT::kTotalLagrangian),
      fLHS(fCtx, opt.linear_tolerance > 0.0 ? opt.linear_tolerance : deck.linear_tolerance, deck.max_linear_iterations)
{
    const char* caller = "FEManagerHIP";
    fElements.SetGeometry(deck.geom.X, deck.conn, deck.elem_block);
    for (size_t m = 0; m < deck.materials.size(); m++)
        fElements.SetMaterial((int)m, deck.materials[m].first, deck.materials[m].second);
    fElements.SetPrecision(opt.precision);
    fElements.SetMaterialModel(deck.material_model);
    if (deck.solver == "newton_krylov_solver")
        fCtx.Check(tg_set_linear_solver(fCtx.Handle(), TG_LIN_GMRES, deck.gmres_restart), caller);
    // one schedule drives every scheduled card: value_k * schedule(t)
    std::vector<int> nodes, dofs;
    std::vector<double> vals;
    for (const KBCCard& k : deck.kbc) {
        if (k.schedule >= 0) {
            if (fSchedule >= 0 && k.schedule != fSchedule)
                ExceptionT::BadInputValue(caller, "kinematic cards on more than one schedule");
            fSchedule = k.schedule;
        }
        nodes.push_back(k.node);
        dofs.push_back(k.dof);
        vals.push_back(k.schedule >= 0 ? k.value : 0.0);
    }
    fElements.SetKinematicBC(nodes, dofs, vals);
    std::vector<int> fn, fd;
    std::vector<double> fv;
    for (const FBCCard& f : deck.fbc) {
        fn.push_back(f.node);
        fd.push_back(f.dof);
        fv.push_back(0.0);
    }
    fElements.SetForces(fn, fd, fv);
    const int n_eq = fElements.InitializeEquations();
    fLHS.Initialize(n_eq, n_eq, 1);
}

const std::vector<StepRecord>& FEManagerHIP::Solve(std::ostream* log)
{
    NLSolverHIP::Parameters p;
    p.max_iterations = fDeck.max_iterations;
    p.min_iterations = fDeck.min_iterations;
    p.abs_tolerance = fDeck.abs_tolerance;
    p.rel_tolerance = fDeck.rel_tolerance;
    p.divergence_tolerance = fDeck.divergence_tolerance;
    p.linear_tolerance = fOpt.linear_tolerance > 0.0 ? fOpt.linear_tolerance : fDeck.linear_tolerance;
    p.max_linear_iterations = fDeck.max_linear_iterations;
    if (fDeck.solver == "PCG_solver") {
        p.method = NLSolverHIP::kNonlinearCG;
        p.restart = fDeck.cg_restart;
        p.search_iterations = fDeck.cg_search_iterations;
        p.orthog_tolerance = fDeck.cg_orthog_tolerance;
        p.max_step = fDeck.cg_max_step;
    }
    NLSolverHIP solver(fElements, fLHS, p, fOpt.driver);
    fSteps.clear();
    for (int n = 1; n <= fDeck.num_steps; n++) {
        const double t = n * fDeck.time_step;
        const double s = fSchedule >= 0 ? fDeck.schedules[(size_t)fSchedule].Value(t) : 0.0;
        fElements.InitStep(s);
        if (!fDeck.fbc.empty()) {
            std::vector<double> f;  // FBC_CardT::CurrentValue
            for (const FBCCard& c : fDeck.fbc)
                f.push_back(c.schedule >= 0 ? c.value * fDeck.schedules[(size_t)c.schedule].Value(t) : c.value);
            fElements.ApplyForces(f);
        }
        const NLSolverHIP::SolutionStatusT st = solver.Solve(-1);
        if (st != NLSolverHIP::kConverged) {
            fElements.ResetStep();
            ExceptionT::GeneralFail("FEManagerHIP::Solve", "no convergence at step " + std::to_string(n));
        }
        fElements.CloseStep();
        StepRecord r{n, t, solver.IterationNumber() + 1, solver.LinearIterations(), solver.Error0(),
                     solver.LastError(), fElements.Displacement()};
        if (log)
            *log << " step " << n << " time " << t << ": " << r.iterations << " Newton, " << r.linear_iterations
                 << " PCG, |R0| " << r.error0 << " |R| " << r.error << std::endl;
        fSteps.push_back(std::move(r));
    }
    return fSteps;
}

void FEManagerHIP::WriteRun(std::ostream& out, int digits) const
{
    const int nnd = fDeck.geom.NumNodes();
    char line[160];
    out << "\n S U M M A R Y :\n\n"
        << " Group number. . . . . . . . . . . . . . . . . . = 0\n"
        << " Number of nodal values. . . . . . . . . . . . . = 3\n"
        << " Labels:\n\tD_X\n\tD_Y\n\tD_Z\n"
        << " Element group . . . . . . . . . . . . . . . . . = " << fDeck.element << "_hip\n\n"
        << " O U T P U T   D A T A :\n";
    for (size_t k = 0; k < fSteps.size(); k++) {
        const StepRecord& r = fSteps[k];
        std::snprintf(line, sizeof line, "%.12e", r.time);
        out << "\n Print Step. . . . . . . . . . . . . . . . . . . = " << k << "\n"
            << " Time. . . . . . . . . . . . . . . . . . . . . . = " << line << "\n"
            << " Newton iterations . . . . . . . . . . . . . . . = " << r.iterations << "\n\n"
            << " Nodal data:\n"
            << " Number of nodal points. . . . . . . . . . . . . = " << nnd << "\n"
            << " Number of values. . . . . . . . . . . . . . . . = 3\n\n"
            << "   index    node                  D_X                  D_Y                  D_Z\n";
        for (int a = 0; a < nnd; a++) {
            const int w = digits + 9;
            std::snprintf(line, sizeof line, "%8d%8d%*.*e%*.*e%*.*e\n", a + 1, a + 1, w, digits, r.u[3L * a], w,
                          digits, r.u[3L * a + 1], w, digits, r.u[3L * a + 2]);
            out << line;
        }
    }
}

}  // namespace tahoe_hip
