// InputDeck.h -- the subset of a Tahoe XML input and its TahoeII geometry
// that the continuum-solid hot path needs.
//
// XML (parameter tree of tahoe/src/main/FEManagerT.cpp and the classes it
// builds; names as in the reference's input decks):
//   <tahoe geometry_file="...">
//     <time num_steps time_step> <schedule_function><piecewise_linear>
//         <OrderedPair x y/>...                        (ScheduleT, one or more)
//     <nodes><field field_name="displacement">
//         <kinematic_BC dof node_ID [type="u"|"fixed"] [schedule] [value]/>
//         <force_BC dof node_ID [schedule] value/>        (FBC_CardT)
//         <mapped_nodes schedule> <Matrix_3x3 A_i_j.../> <mapped_node_ID_list>
//                                                      (MappedPeriodicT without
//                                                       leader/follower pairs)
//     <element_list><total_lagrangian | updated_lagrangian |
//                    updated_lagrangian_Q1P0 (also *_hip)>
//         <hexahedron [num_ip="8"]/>
//         <large_strain_element_block><block_ID_list><String value/>
//             <large_strain_material_3D><Simo_isotropic | large_strain_StVenant>
//                 <E_and_nu Young_modulus Poisson_ratio/>
//     <nonlinear_solver rel_tolerance abs_tolerance divergence_tolerance
//                       max_iterations [min_iterations]>
//       or <newton_krylov_solver ... [gmres_restart] [linear_tolerance]
//                                    [max_linear_iter]>       (NLSolver_NK)
//       or <PCG_solver ... restart [line_search_iterations]
//                          [line_search_tolerance] [max_step]> (PCGSolver_LS)
//         <any matrix: SPOOLES_matrix, profile_matrix, ...,
//          HIP_MSR_matrix [linear_tolerance] [max_linear_iterations]>
// Anything else on the path (other elements, materials, BC controllers,
// more than one field) is rejected with kBadInputValue naming the element.
//
// Geometry: TahoeII text format (toolbox/src/dataio/input/TahoeInputT.cpp,
// ModelFileT): *dimensions, *nodesets, *sidesets, *elements, *nodes with
// '#' comments; node and element numbers are 1-based in the file.
#ifndef TAHOE_HIP_INPUT_DECK_H
#define TAHOE_HIP_INPUT_DECK_H

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace tahoe_hip {

// minimal XML element tree: name, attributes, children
struct XMLNode {
    std::string name;
    std::map<std::string, std::string> attr;
    std::vector<std::unique_ptr<XMLNode>> children;

    const XMLNode* Child(const std::string& n) const;
    std::vector<const XMLNode*> Children(const std::string& n) const;
    bool Has(const std::string& a) const { return attr.count(a) != 0; }
    std::string Attr(const std::string& a, const std::string& dflt = "") const;
    double Real(const std::string& a, double dflt) const;
    int Int(const std::string& a, int dflt) const;
};

std::unique_ptr<XMLNode> ParseXML(const std::string& text, const std::string& source);

struct Geometry {
    int nsd = 3;
    std::vector<double> X;                         // [nnd][3]
    std::vector<int> block_ids;                    // element set IDs in file order
    std::vector<std::vector<int>> block_conn;      // per set, [nel][8] 0-based
    std::map<int, std::vector<int>> node_sets;     // ID -> 0-based nodes
    int NumNodes() const { return (int)(X.size() / 3); }
};

Geometry ReadGeometry(const std::string& path);

struct ScheduleT {
    std::vector<std::pair<double, double>> points;  // piecewise linear (x, y)
    double Value(double t) const;
};

struct KBCCard {
    int node, dof;       // 0-based
    double value;        // multiplies the schedule value
    int schedule;        // 0-based schedule, -1 for fixed
};

struct FBCCard {
    int node, dof;       // 0-based
    double value;        // multiplies the schedule value
    int schedule;        // 0-based schedule, -1 for none (constant value)
};

struct InputDeck {
    std::string xml_path, geometry_path, title;
    int num_steps = 1;
    double time_step = 1.0;
    std::vector<ScheduleT> schedules;
    std::string element;                 // "total_lagrangian" / "updated_lagrangian" / "updated_lagrangian_Q1P0"
    int formulation = 0;                 // TG_TOTAL_LAGRANGIAN / TG_Q1P0
    Geometry geom;
    std::vector<int> conn;               // all element blocks [nel][8] 0-based
    std::vector<int> elem_block;         // material index of every element
    std::vector<std::pair<double, double>> materials;  // (E, nu) per material index
    std::vector<KBCCard> kbc;
    std::vector<FBCCard> fbc;
    int material_model = 0;              // TG_MAT_SIMO_ISOTROPIC / TG_MAT_STVENANT
    // nonlinear_solver
    double rel_tolerance = 1e-10, abs_tolerance = 1e-14, divergence_tolerance = 10.0;
    int max_iterations = 20, min_iterations = 0;
    double linear_tolerance = 1e-10;
    int max_linear_iterations = 100000;
    std::string matrix;                  // matrix element named in the deck
    std::string solver = "nonlinear_solver";  // or "newton_krylov_solver" (GMRES inner solve), "PCG_solver"
    int gmres_restart = 30;
    // PCG_solver (nonlinear CG with line search)
    int cg_restart = 0, cg_search_iterations = 3;
    double cg_orthog_tolerance = 0.25, cg_max_step = 2.5;
};

// reads the XML and the geometry it names (relative to the XML's folder)
InputDeck ReadInputDeck(const std::string& xml_path);

}  // namespace tahoe_hip

#endif
