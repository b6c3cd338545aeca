// InputDeck.cpp -- XML subset, TahoeII geometry and the deck they describe.
#include "InputDeck.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "ExceptionT.h"
#include "tahoe_gpu.h"

namespace tahoe_hip {

// ---------------------------------------------------------------------------
// XML
// ---------------------------------------------------------------------------

const XMLNode* XMLNode::Child(const std::string& n) const
{
    for (const auto& c : children)
        if (c->name == n) return c.get();
    return nullptr;
}

std::vector<const XMLNode*> XMLNode::Children(const std::string& n) const
{
    std::vector<const XMLNode*> out;
    for (const auto& c : children)
        if (c->name == n) out.push_back(c.get());
    return out;
}

std::string XMLNode::Attr(const std::string& a, const std::string& dflt) const
{
    auto it = attr.find(a);
    return it == attr.end() ? dflt : it->second;
}

static double to_real(const std::string& s, const std::string& what)
{
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (s.empty() || (end && *end != '\0' && !std::isspace((unsigned char)*end)))
        ExceptionT::BadInputValue("InputDeck", "not a number for " + what + ": \"" + s + "\"");
    return v;
}

static int to_int(const std::string& s, const std::string& what)
{
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (s.empty() || (end && *end != '\0' && !std::isspace((unsigned char)*end)))
        ExceptionT::BadInputValue("InputDeck", "not an integer for " + what + ": \"" + s + "\"");
    return (int)v;
}

double XMLNode::Real(const std::string& a, double dflt) const { return Has(a) ? to_real(Attr(a), name + "." + a) : dflt; }
int XMLNode::Int(const std::string& a, int dflt) const { return Has(a) ? to_int(Attr(a), name + "." + a) : dflt; }

namespace {

struct XMLReader {
    const std::string& s;
    const std::string& src;
    size_t i = 0;

    [[noreturn]] void error(const std::string& msg)
    {
        const long line = 1 + std::count(s.begin(), s.begin() + (long)std::min(i, s.size()), '\n');
        ExceptionT::BadInputValue("ParseXML", src + ":" + std::to_string(line) + ": " + msg);
    }
    void skip_ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool starts(const char* t) const { return s.compare(i, std::char_traits<char>::length(t), t) == 0; }
    void skip_misc()
    {
        // text, comments, declarations and processing instructions between elements
        for (;;) {
            while (i < s.size() && s[i] != '<') i++;
            if (starts("<!--")) {
                size_t e = s.find("-->", i + 4);
                if (e == std::string::npos) error("unterminated comment");
                i = e + 3;
            } else if (starts("<?")) {
                size_t e = s.find("?>", i + 2);
                if (e == std::string::npos) error("unterminated declaration");
                i = e + 2;
            } else if (starts("<!")) {
                size_t e = s.find('>', i + 2);
                if (e == std::string::npos) error("unterminated declaration");
                i = e + 1;
            } else
                return;
        }
    }
    std::string name()
    {
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == ':' || s[i] == '-' ||
                                s[i] == '.'))
            i++;
        if (i == b) error("expecting a name");
        return s.substr(b, i - b);
    }
    static std::string unescape(const std::string& v)
    {
        std::string o;
        for (size_t k = 0; k < v.size(); k++) {
            if (v[k] == '&') {
                const size_t e = v.find(';', k);
                const std::string ent = e == std::string::npos ? "" : v.substr(k + 1, e - k - 1);
                const char* rep = ent == "lt" ? "<" : ent == "gt" ? ">" : ent == "amp" ? "&" : ent == "quot" ? "\"" :
                                  ent == "apos" ? "'" : nullptr;
                if (rep) {
                    o += rep;
                    k = e;
                    continue;
                }
            }
            o += v[k];
        }
        return o;
    }
    std::unique_ptr<XMLNode> element()
    {
        if (i >= s.size() || s[i] != '<') error("expecting an element");
        i++;
        auto n = std::make_unique<XMLNode>();
        n->name = name();
        for (;;) {
            skip_ws();
            if (i >= s.size()) error("unterminated element <" + n->name + ">");
            if (starts("/>")) {
                i += 2;
                return n;
            }
            if (s[i] == '>') {
                i++;
                break;
            }
            const std::string a = name();
            skip_ws();
            if (i >= s.size() || s[i] != '=') error("expecting '=' after attribute " + a);
            i++;
            skip_ws();
            if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) error("expecting a quoted value for " + a);
            const char q = s[i++];
            const size_t e = s.find(q, i);
            if (e == std::string::npos) error("unterminated value of " + a);
            n->attr[a] = unescape(s.substr(i, e - i));
            i = e + 1;
        }
        for (;;) {
            skip_misc();
            if (i >= s.size()) error("missing </" + n->name + ">");
            if (starts("</")) {
                i += 2;
                const std::string close = name();
                if (close != n->name) error("</" + close + "> closes <" + n->name + ">");
                skip_ws();
                if (i >= s.size() || s[i] != '>') error("expecting '>'");
                i++;
                return n;
            }
            n->children.push_back(element());
        }
    }
};

}  // namespace

std::unique_ptr<XMLNode> ParseXML(const std::string& text, const std::string& source)
{
    XMLReader r{text, source};
    r.skip_misc();
    auto root = r.element();
    r.skip_misc();
    if (r.i < text.size()) r.error("content after the root element");
    return root;
}

// ---------------------------------------------------------------------------
// TahoeII geometry
// ---------------------------------------------------------------------------

namespace {

struct Tokens {
    std::vector<std::string> t;
    size_t i = 0;
    std::string path;

    explicit Tokens(const std::string& p) : path(p)
    {
        std::ifstream in(p);
        if (!in) ExceptionT::Throw(ExceptionT::kDatabaseFail, "ReadGeometry", "cannot open " + p);
        std::string line;
        while (std::getline(in, line)) {
            const size_t h = line.find('#');
            if (h != std::string::npos) line.resize(h);
            std::istringstream ss(line);
            std::string w;
            while (ss >> w) t.push_back(w);
        }
    }
    [[noreturn]] void error(const std::string& msg) const
    {
        ExceptionT::Throw(ExceptionT::kDatabaseFail, "ReadGeometry", path + ": " + msg);
    }
    const std::string& next()
    {
        if (i >= t.size()) error("unexpected end of file");
        return t[i++];
    }
    int integer() { return to_int(next(), "geometry"); }
    double real() { return to_real(next(), "geometry"); }
    void expect(const std::string& key)
    {
        const std::string& w = next();
        if (w != key) error("expecting " + key + ", found " + w);
    }
    bool seek(const std::string& key)
    {
        for (size_t k = i; k < t.size(); k++)
            if (t[k] == key) {
                i = k + 1;
                return true;
            }
        return false;
    }
};

}  // namespace

Geometry ReadGeometry(const std::string& path)
{
    Tokens tk(path);
    Geometry g;
    if (!tk.seek("*dimensions")) tk.error("no *dimensions");
    const int nnd = tk.integer();
    g.nsd = tk.integer();
    if (g.nsd != 3) tk.error("only 3D geometry is on the path");
    const int nblk = tk.integer();
    std::vector<int> blk_nel(nblk), blk_nen(nblk);
    for (int b = 0; b < nblk; b++) {
        g.block_ids.push_back(tk.integer());
        blk_nel[b] = tk.integer();
        blk_nen[b] = tk.integer();
    }
    const int nsets = tk.integer();
    std::vector<int> set_ids(nsets), set_n(nsets);
    for (int k = 0; k < nsets; k++) {
        set_ids[k] = tk.integer();
        set_n[k] = tk.integer();
    }
    if (!tk.seek("*nodesets")) tk.error("no *nodesets");
    for (int k = 0; k < nsets; k++) {
        tk.expect("*set");
        const int n = tk.integer();
        if (n != set_n[k]) tk.error("node set size differs from *dimensions");
        std::vector<int>& v = g.node_sets[set_ids[k]];
        for (int j = 0; j < n; j++) {
            const int node = tk.integer();
            if (n == 1 && node == -1) {  // a one-entry set "-1" holds every node
                for (int a = 0; a < nnd; a++) v.push_back(a);
                break;
            }
            if (node < 1 || node > nnd) tk.error("node set refers to a missing node");
            v.push_back(node - 1);
        }
    }
    if (!tk.seek("*elements")) tk.error("no *elements");
    g.block_conn.resize(nblk);
    for (int b = 0; b < nblk; b++) {
        tk.expect("*set");
        const int nel = tk.integer(), nen = tk.integer();
        if (nel != blk_nel[b] || nen != blk_nen[b]) tk.error("element set size differs from *dimensions");
        if (nen != 8) tk.error("only 8-node hexahedra are on the path");
        std::vector<int>& c = g.block_conn[b];
        c.reserve((size_t)nel * 8);
        for (int e = 0; e < nel; e++) {
            (void)tk.integer();  // element number
            for (int a = 0; a < 8; a++) {
                const int node = tk.integer();
                if (node < 1 || node > nnd) tk.error("element refers to a missing node");
                c.push_back(node - 1);
            }
        }
    }
    if (!tk.seek("*nodes")) tk.error("no *nodes");
    if (tk.integer() != nnd || tk.integer() != 3) tk.error("*nodes header differs from *dimensions");
    g.X.assign((size_t)nnd * 3, 0.0);
    for (int a = 0; a < nnd; a++) {
        const int id = tk.integer();
        if (id != a + 1) tk.error("nodes must be numbered 1..nnd in order");
        for (int c = 0; c < 3; c++) g.X[3L * a + c] = tk.real();
    }
    return g;
}

// ---------------------------------------------------------------------------
// deck
// ---------------------------------------------------------------------------

double ScheduleT::Value(double t) const
{
    if (points.empty()) return 0.0;
    if (t <= points.front().first) return points.front().second;
    if (t >= points.back().first) return points.back().second;
    for (size_t k = 1; k < points.size(); k++)
        if (t < points[k].first) {
            const auto& a = points[k - 1];
            const auto& b = points[k];
            return a.second + (t - a.first) * (b.second - a.second) / (b.first - a.first);
        }
    return points.back().second;
}

static std::string folder_of(const std::string& p)
{
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

static void reject_unknown(const XMLNode& n, const std::vector<std::string>& known)
{
    for (const auto& c : n.children)
        if (std::find(known.begin(), known.end(), c->name) == known.end())
            ExceptionT::BadInputValue("ReadInputDeck", "<" + c->name + "> inside <" + n.name + "> is not supported");
}

static std::vector<int> string_list(const XMLNode* list)
{
    std::vector<int> ids;
    if (!list) return ids;
    for (const XMLNode* s : list->Children("String")) ids.push_back(to_int(s->Attr("value"), "ID list"));
    return ids;
}

InputDeck ReadInputDeck(const std::string& xml_path)
{
    const char* caller = "ReadInputDeck";
    std::ifstream in(xml_path);
    if (!in) ExceptionT::Throw(ExceptionT::kDatabaseFail, caller, "cannot open " + xml_path);
    std::stringstream buf;
    buf << in.rdbuf();
    auto root = ParseXML(buf.str(), xml_path);
    if (root->name != "tahoe") ExceptionT::BadInputValue(caller, "root element must be <tahoe>");
    reject_unknown(*root, {"time", "nodes", "element_list", "nonlinear_solver", "newton_krylov_solver", "PCG_solver"});

    InputDeck d;
    d.xml_path = xml_path;
    d.title = root->Attr("title");
    const std::string gf = root->Attr("geometry_file");
    if (gf.empty()) ExceptionT::BadInputValue(caller, "no geometry_file");
    d.geometry_path = gf[0] == '/' ? gf : folder_of(xml_path) + "/" + gf;
    d.geom = ReadGeometry(d.geometry_path);

    // ---- time and schedules (TimeManagerT) ----
    const XMLNode* time = root->Child("time");
    if (!time) ExceptionT::BadInputValue(caller, "no <time>");
    d.num_steps = time->Int("num_steps", 1);
    d.time_step = time->Real("time_step", 1.0);
    if (d.num_steps < 0 || !(d.time_step > 0.0)) ExceptionT::BadInputValue(caller, "num_steps and time_step");
    for (const XMLNode* sf : time->Children("schedule_function")) {
        const XMLNode* pl = sf->Child("piecewise_linear");
        if (!pl || sf->children.size() != 1) ExceptionT::BadInputValue(caller, "only piecewise_linear schedules");
        ScheduleT s;
        for (const XMLNode* p : pl->Children("OrderedPair")) s.points.emplace_back(p->Real("x", 0.0), p->Real("y", 0.0));
        if (s.points.empty()) ExceptionT::BadInputValue(caller, "empty schedule");
        d.schedules.push_back(s);
    }

    // ---- boundary conditions (FieldT, KBC_CardT, MappedPeriodicT) ----
    const XMLNode* nodes = root->Child("nodes");
    if (!nodes) ExceptionT::BadInputValue(caller, "no <nodes>");
    reject_unknown(*nodes, {"field"});
    const auto fields = nodes->Children("field");
    if (fields.size() != 1) ExceptionT::BadInputValue(caller, "exactly one field (displacement)");
    const XMLNode& field = *fields[0];
    reject_unknown(field, {"dof_labels", "kinematic_BC", "force_BC", "mapped_nodes"});
    auto node_set = [&](const std::string& id) -> const std::vector<int>& {
        auto it = d.geom.node_sets.find(to_int(id, "node_ID"));
        if (it == d.geom.node_sets.end()) ExceptionT::BadInputValue(caller, "no node set " + id);
        return it->second;
    };
    auto schedule_index = [&](int one_based) {
        if (one_based < 1 || one_based > (int)d.schedules.size())
            ExceptionT::BadInputValue(caller, "schedule " + std::to_string(one_based) + " is not defined");
        return one_based - 1;
    };
    for (const XMLNode* k : field.Children("kinematic_BC")) {
        const int dof = k->Int("dof", 0) - 1;
        if (dof < 0 || dof > 2) ExceptionT::BadInputValue(caller, "kinematic_BC dof must be 1..3");
        const std::string type = k->Attr("type", "fixed");
        double value = 0.0;
        int sched = -1;
        if (type == "u" || type == "D_X" || type == "D_Y" || type == "D_Z") {
            value = k->Real("value", 0.0);
            sched = k->Has("schedule") ? schedule_index(k->Int("schedule", 0)) : -1;
            if (sched < 0) value = 0.0;
        } else if (type != "fixed")
            ExceptionT::BadInputValue(caller, "kinematic_BC type \"" + type + "\" is not on the path");
        for (int n : node_set(k->Attr("node_ID"))) d.kbc.push_back({n, dof, value, sched});
    }
    // FieldT's force cards (FieldT.cpp:1370, FBC_CardT::SetValues): value
    // times the schedule's value, or the value itself without a schedule
    for (const XMLNode* f : field.Children("force_BC")) {
        const int dof = f->Int("dof", 0) - 1;
        if (dof < 0 || dof > 2) ExceptionT::BadInputValue(caller, "force_BC dof must be 1..3");
        const double value = f->Real("value", 0.0);
        const int sched = f->Has("schedule") ? schedule_index(f->Int("schedule", 0)) : -1;
        for (int n : node_set(f->Attr("node_ID"))) d.fbc.push_back({n, dof, value, sched});
    }
    for (const XMLNode* m : field.Children("mapped_nodes")) {
        reject_unknown(*m, {"Matrix_3x3", "mapped_node_ID_list"});
        const int sched = schedule_index(m->Int("schedule", 0));
        double A[3][3] = {{0}};
        if (const XMLNode* mx = m->Child("Matrix_3x3"))
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    A[i][j] = mx->Real("A_" + std::to_string(i + 1) + "_" + std::to_string(j + 1), 0.0);
        for (int id : string_list(m->Child("mapped_node_ID_list")))
            for (int n : node_set(std::to_string(id))) {
                const double* X = &d.geom.X[3L * n];
                for (int j = 0; j < 3; j++)
                    d.kbc.push_back({n, j, A[j][0] * X[0] + A[j][1] * X[1] + A[j][2] * X[2], sched});
            }
    }

    // ---- element group (ElementListT) ----
    const XMLNode* el = root->Child("element_list");
    if (!el || el->children.size() != 1) ExceptionT::BadInputValue(caller, "exactly one element group");
    const XMLNode& grp = *el->children[0];
    std::string ename = grp.name;
    if (ename.size() > 4 && ename.compare(ename.size() - 4, 4, "_hip") == 0) ename.resize(ename.size() - 4);
    // updated_lagrangian integrates the same forms over the current
    // configuration (UpdatedLagrangianT.cpp:83-92); it runs on the
    // total-Lagrangian kernels
    if (ename == "total_lagrangian" || ename == "updated_lagrangian")
        d.formulation = TG_TOTAL_LAGRANGIAN;
    else if (ename == "updated_lagrangian_Q1P0")
        d.formulation = TG_Q1P0;
    else
        ExceptionT::BadInputValue(caller, "element group <" + grp.name + "> is not on the path");
    d.element = ename;
    reject_unknown(grp, {"hexahedron", "solid_element_nodal_output", "solid_element_element_output",
                         "large_strain_element_block"});
    const XMLNode* hex = grp.Child("hexahedron");
    if (!hex) ExceptionT::BadInputValue(caller, "only <hexahedron> geometry is on the path");
    if (hex->Int("num_ip", 8) != 8) ExceptionT::BadInputValue(caller, "hexahedron num_ip must be 8");
    std::map<int, int> block_material;  // element set ID -> material index
    int model = -1;
    for (const XMLNode* b : grp.Children("large_strain_element_block")) {
        const XMLNode* m3 = b->Child("large_strain_material_3D");
        const XMLNode* mat = nullptr;
        int this_model = -1;
        if (m3 && (mat = m3->Child("Simo_isotropic")))
            this_model = TG_MAT_SIMO_ISOTROPIC;
        else if (m3 && (mat = m3->Child("large_strain_StVenant")))
            this_model = TG_MAT_STVENANT;
        else
            ExceptionT::BadInputValue(caller, "only Simo_isotropic and large_strain_StVenant materials are on the path");
        if (model >= 0 && this_model != model)
            ExceptionT::BadInputValue(caller, "all blocks of the group must use one material model");
        model = this_model;
        const XMLNode* en = mat->Child("E_and_nu");
        if (!en) ExceptionT::BadInputValue(caller, mat->name + " needs <E_and_nu>");
        const int mid = (int)d.materials.size();
        d.materials.emplace_back(en->Real("Young_modulus", 0.0), en->Real("Poisson_ratio", 0.0));
        for (int id : string_list(b->Child("block_ID_list"))) block_material[id] = mid;
    }
    for (size_t b = 0; b < d.geom.block_ids.size(); b++) {
        auto it = block_material.find(d.geom.block_ids[b]);
        if (it == block_material.end()) continue;  // element set not in this group
        const std::vector<int>& c = d.geom.block_conn[b];
        d.conn.insert(d.conn.end(), c.begin(), c.end());
        d.elem_block.insert(d.elem_block.end(), c.size() / 8, it->second);
    }
    if (d.conn.empty()) ExceptionT::BadInputValue(caller, "the element group has no elements");
    d.material_model = model < 0 ? TG_MAT_SIMO_ISOTROPIC : model;

    // ---- nonlinear solver (NLSolver, or NLSolver_NK: Newton with GMRES) ----
    const XMLNode* ns = root->Child("nonlinear_solver");
    const XMLNode* nk = root->Child("newton_krylov_solver");
    const XMLNode* cg = root->Child("PCG_solver");
    if ((ns != nullptr) + (nk != nullptr) + (cg != nullptr) > 1) ExceptionT::BadInputValue(caller, "one solver per deck");
    if (cg) {
        // PCGSolver_LS parameters (PCGSolver_LS.cpp:48-98)
        d.solver = "PCG_solver";
        d.cg_restart = cg->Int("restart", 0);
        d.cg_search_iterations = cg->Int("line_search_iterations", 3);
        d.cg_orthog_tolerance = cg->Real("line_search_tolerance", 0.25);
        d.cg_max_step = cg->Real("max_step", 2.5);
        if (d.cg_restart < 0 || d.cg_search_iterations < 0)
            ExceptionT::BadInputValue(caller, "PCG_solver restart and line_search_iterations must be >= 0");
        ns = cg;
    }
    if (nk) {
        // NLSolver_NK parameters (NLSolver_NK.cpp:17-63): gmres_restart 30,
        // linear_tolerance 1e-3, max_linear_iter 0 (= 3 n)
        d.solver = "newton_krylov_solver";
        d.gmres_restart = nk->Int("gmres_restart", 30);
        if (d.gmres_restart < 1) ExceptionT::BadInputValue(caller, "gmres_restart must be >= 1");
        d.linear_tolerance = nk->Real("linear_tolerance", 1.0e-3);
        d.max_linear_iterations = nk->Int("max_linear_iter", 0);
        ns = nk;
    }
    if (!ns)
        ExceptionT::BadInputValue(caller,
                                  "only <nonlinear_solver>, <newton_krylov_solver> and <PCG_solver> are on the path");
    d.rel_tolerance = ns->Real("rel_tolerance", d.rel_tolerance);
    d.abs_tolerance = ns->Real("abs_tolerance", d.abs_tolerance);
    d.divergence_tolerance = ns->Real("divergence_tolerance", d.divergence_tolerance);
    d.max_iterations = ns->Int("max_iterations", d.max_iterations);
    d.min_iterations = ns->Int("min_iterations", d.min_iterations);
    if (ns->children.size() == 1) {
        const XMLNode& m = *ns->children[0];
        d.matrix = m.name;
        if (m.name == "HIP_MSR_matrix" && !nk) {
            d.linear_tolerance = m.Real("linear_tolerance", d.linear_tolerance);
            d.max_linear_iterations = m.Int("max_linear_iterations", d.max_linear_iterations);
        }
    } else if (ns->children.size() > 1)
        ExceptionT::BadInputValue(caller, "one matrix inside <nonlinear_solver>");
    return d;
}

}  // namespace tahoe_hip
