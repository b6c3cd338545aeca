// tahoe_hip -- run a Tahoe XML input deck of the continuum-solid hot path on
// one MI355X.
//
//   tahoe_hip -f deck.xml [-o out.run] [-precision fp64|mixed]
//             [-driver device|host] [-device N] [-linear_tolerance x]
//             [-digits n] [-check]
//
// -f        the input deck (tahoe -f, tahoe/src/main/main.cpp); its
//           geometry_file is read relative to the deck
// -o        nodal displacements of every step, laid out like <root>.io0.run
//           (default: <deck without .xml>.io0.run)
// -driver   device: Newton loop resident on the GPU; host: the reference's
//           NLSolver loop over the element group and matrix plugins
// -check    read and check the deck only (no GPU)
// Exit status: 0, or the ExceptionT code of the failure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "FEManagerHIP.h"

using namespace tahoe_hip;

static int usage()
{
    std::cerr << "usage: tahoe_hip -f deck.xml [-o out.run] [-precision fp64|mixed] [-driver device|host]\n"
                 "                 [-device N] [-linear_tolerance x] [-digits n] [-check]\n";
    return ExceptionT::kBadInputValue;
}

int main(int argc, char** argv)
{
    std::string deck_path, out_path;
    FEManagerHIP::Options opt;
    bool check = false;
    int digits = 12;
    for (int k = 1; k < argc; k++) {
        const std::string a = argv[k];
        auto value = [&]() -> std::string {
            if (k + 1 >= argc) {
                std::cerr << a << " needs a value\n";
                std::exit(usage());
            }
            return argv[++k];
        };
        if (a == "-f")
            deck_path = value();
        else if (a == "-o")
            out_path = value();
        else if (a == "-precision") {
            const std::string v = value();
            if (v == "fp64")
                opt.precision = TG_FP64;
            else if (v == "mixed")
                opt.precision = TG_MIXED;
            else
                return usage();
        } else if (a == "-driver") {
            const std::string v = value();
            if (v == "device")
                opt.driver = NLSolverHIP::kDevice;
            else if (v == "host")
                opt.driver = NLSolverHIP::kHost;
            else
                return usage();
        } else if (a == "-device")
            opt.device = std::atoi(value().c_str());
        else if (a == "-linear_tolerance")
            opt.linear_tolerance = std::atof(value().c_str());
        else if (a == "-digits")
            digits = std::atoi(value().c_str());
        else if (a == "-check")
            check = true;
        else
            return usage();
    }
    if (deck_path.empty() || digits < 1 || digits > 17) return usage();
    if (out_path.empty()) {
        out_path = deck_path;
        if (out_path.size() > 4 && out_path.compare(out_path.size() - 4, 4, ".xml") == 0)
            out_path.resize(out_path.size() - 4);
        out_path += ".io0.run";
    }
    try {
        const InputDeck deck = ReadInputDeck(deck_path);
        int scheduled = 0;
        for (const KBCCard& c : deck.kbc) scheduled += c.schedule >= 0;
        std::cout << " deck " << deck_path << ": " << deck.element << ", " << deck.geom.NumNodes() << " nodes, "
                  << deck.conn.size() / 8 << " elements, " << deck.materials.size() << " materials, "
                  << deck.kbc.size() << " kinematic cards (" << scheduled << " scheduled), " << deck.fbc.size()
                  << " force cards, " << (deck.material_model == TG_MAT_STVENANT ? "large_strain_StVenant"
                                                                                  : "Simo_isotropic")
                  << ", " << deck.num_steps << " steps, matrix " << (deck.matrix.empty() ? "-" : deck.matrix)
                  << std::endl;
        if (check) return 0;
        FEManagerHIP fe(deck, opt);
        std::cout << " equations " << fe.Elements().NumEquations() << ", MSR length " << fe.Matrix().MSRLength()
                  << std::endl;
        fe.Solve(&std::cout);
        std::ofstream out(out_path);
        if (!out) ExceptionT::Throw(ExceptionT::kDatabaseFail, "tahoe_hip", "cannot write " + out_path);
        fe.WriteRun(out, digits);
        std::cout << " End Execution" << std::endl;
        return 0;
    } catch (const ExceptionT& e) {
        std::cerr << " tahoe_hip: " << e.what() << std::endl;
        return (int)e.Code();
    }
}
