// dmpci.cpp -- reader of Osprey-DPD control files and the polymer topology they define.
//
// Follows the keyword order of CInputDPDFile::Serialize (src/InputDPDFile.cpp:327-1052):
//   dpd  Title " .. "  Date d  Comment " .. "  State <type> <data>  [Wall|Gravity ..]
//   Bead blocks  Bond*  BondPair*  Polymer*  Box  Density  Temp  RNGSeed  Lambda  Step
//   Time  SamplePeriod  AnalysisPeriod  DensityPeriod  DisplayPeriod  RestartPeriod  Grid
//   then Process / Analysis / Event / Command blocks.
// Only the DPD subset that drives the step loop is interpreted; state-type data other
// than restart ids, and Process/Analysis/Event blocks, are read past.
#include "dmpci.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <set>
#include <sstream>

namespace dpd {
namespace host {

namespace {

const char* kNameChars = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789-_";

bool is_name(const std::string& s)
{
    // CInputData::IsNameValid: valid characters, not starting with a digit
    return !s.empty() && s.find_first_not_of(kNameChars) == std::string::npos && !(s[0] >= '0' && s[0] <= '9');
}

bool is_number(const std::string& s) { return !s.empty() && s.find_first_not_of("0123456789") == std::string::npos; }

struct Tokens {
    std::istringstream in;
    explicit Tokens(const std::string& t) : in(t) {}
    bool next(std::string& tok) { return (bool)(in >> tok); }
    template <class T> bool value(T& v) { return (bool)(in >> v); }
};

bool quoted(Tokens& t, std::string& out)
{
    std::string tok;
    if (!t.next(tok) || tok != "\"") return false;
    out.clear();
    while (t.next(tok) && tok != "\"") out += (out.empty() ? "" : " ") + tok;
    return tok == "\"";
}

// ---- polymer shape: S-expression tree ------------------------------------------

struct Node {
    bool leaf = true;
    std::string name;
    std::vector<Node> kids;
};

bool parse_list(const std::string& s, size_t& p, Node& out, std::string& err)
{
    // s[p] == '('
    out.leaf = false;
    ++p;
    while (p < s.size()) {
        const char c = s[p];
        if (c == ' ' || c == '\t' || c == '\n') { ++p; continue; }
        if (c == ')') { ++p; return true; }
        if (c == '(') {
            Node k;
            if (!parse_list(s, p, k, err)) return false;
            out.kids.push_back(k);
            continue;
        }
        size_t q = p;
        while (q < s.size() && s[q] != ' ' && s[q] != '(' && s[q] != ')' && s[q] != '\t' && s[q] != '\n') ++q;
        Node k;
        k.name = s.substr(p, q - p);
        out.kids.push_back(k);
        p = q;
    }
    err = "unbalanced brackets in polymer shape";
    return false;
}

struct Expander {
    const InputData& d;
    const std::map<std::string, int>& bond_names;
    PolymerType& pt;
    std::map<int, std::pair<int, std::string>> loops;   // loop id -> (bead index, name)
    std::string err;

    // CInputData::AddBeadtoPolymer (src/InputData.cpp:1469-1513): the new bead is bonded
    // to the current bead (head = current, tail = new) with the bond type named by the
    // two bead names; the current index moves when `inc`.
    bool add_bead(const std::string& cur_name, const std::string& name, int& cur, bool inc)
    {
        const int t = d.type_of(name);
        if (t < 0) { err = "unknown bead name in polymer shape: " + name; return false; }
        pt.bead_type.push_back(t);
        const int idx = (int)pt.bead_type.size() - 1;
        if (idx > 0) {
            auto it = bond_names.find(cur_name + name);
            if (it == bond_names.end()) { err = "no bond type for " + cur_name + " " + name; return false; }
            pt.bond_head.push_back(cur);
            pt.bond_tail.push_back(idx);
            pt.bond_type.push_back(it->second);
        }
        if (inc) cur = idx;
        return true;
    }

    // sequential elements with a moving current bead (ParseElement's compound case)
    bool seq(const std::vector<Node>& els, size_t from, std::string& cur_name, int& cur)
    {
        for (size_t i = from; i < els.size(); ++i)
            if (!element(els[i], cur_name, cur)) return false;
        return true;
    }

    bool element(const Node& e, std::string& cur_name, int& cur)
    {
        if (e.leaf) {
            if (!is_name(e.name)) { err = "invalid element in polymer shape: " + e.name; return false; }
            if (!add_bead(cur_name, e.name, cur, true)) return false;
            cur_name = e.name;
            return true;
        }
        if (e.kids.empty()) return true;
        const Node& h = e.kids[0];
        if (h.leaf && is_number(h.name)) {
            // (n X Y ..): the remaining elements repeated n times (ExpandElement, :813-833)
            const int n = std::atoi(h.name.c_str());
            for (int r = 0; r < n; ++r)
                if (!seq(e.kids, 1, cur_name, cur)) return false;
            return true;
        }
        if (h.leaf && h.name == "*") {
            // side branch from the current bead; the main chain continues from it afterwards
            if (e.kids.size() == 2 && e.kids[1].leaf && is_name(e.kids[1].name))
                return add_bead(cur_name, e.kids[1].name, cur, false);
            std::string bname = cur_name;
            int bcur = cur;
            return seq(e.kids, 1, bname, bcur);
        }
        if (h.leaf && h.name == "/") {
            // (/ id Name): a bead that closes a loop with the bead first tagged by id
            if (e.kids.size() != 3 || !e.kids[1].leaf || !e.kids[2].leaf) { err = "bad loop element"; return false; }
            const int lid = std::atoi(e.kids[1].name.c_str());
            const std::string& target = e.kids[2].name;
            if (!add_bead(cur_name, target, cur, true)) return false;
            cur_name = target;
            auto it = loops.find(lid);
            if (it == loops.end()) {
                loops[lid] = {cur, target};
            } else {
                auto bt = bond_names.find(cur_name + it->second.second);
                if (bt == bond_names.end()) { err = "no bond type closing loop"; return false; }
                pt.bond_head.push_back(cur);
                pt.bond_tail.push_back(it->second.first);
                pt.bond_type.push_back(bt->second);
            }
            return true;
        }
        return seq(e.kids, 0, cur_name, cur);
    }
};

}  // namespace

std::vector<std::string> InputData::state_values(const std::string& key, size_t count) const
{
    for (size_t i = 0; i < state_tokens.size(); ++i)
        if (state_tokens[i] == key && i + count < state_tokens.size() + 1)
            return std::vector<std::string>(state_tokens.begin() + i + 1,
                                            state_tokens.begin() + std::min(state_tokens.size(), i + 1 + count));
    return {};
}

int InputData::type_of(const std::string& name) const
{
    for (size_t i = 0; i < bead_names.size(); ++i)
        if (bead_names[i] == name) return (int)i;
    return -1;
}

bool read_dmpci(const std::string& text, InputData& d, std::string& err)
{
    d = InputData();
    Tokens t(text);
    std::string tok;
    auto fail = [&](const std::string& m) { err = m; return false; };
    if (!t.next(tok) || tok != "dpd") return fail("Unknown simulation type");
    if (!t.next(tok) || tok != "Title" || !quoted(t, d.title)) return fail("Error reading Title");
    if (!t.next(tok) || tok != "Date" || !t.next(d.date)) return fail("Error reading Date");
    if (!t.next(tok) || tok != "Comment" || !quoted(t, d.comment)) return fail("Error reading Comment");
    if (!t.next(tok) || tok != "State" || !t.next(d.state)) return fail("Error reading State");
    // state data: all tokens up to the first Wall / Gravity / Bead
    while (t.next(tok) && tok != "Bead" && tok != "Wall" && tok != "Gravity") d.state_tokens.push_back(tok);
    if (d.state == "restart") {
        // CInitialStateRestart::get (src/InitialStateRestart.cpp:96-139): RunId <id> StateId <n>
        if (d.state_tokens.size() < 4 || d.state_tokens[0] != "RunId" || d.state_tokens[2] != "StateId")
            return fail("Error reading initial state data");
        d.restart_run = d.state_tokens[1];
        d.restart_state = std::atoll(d.state_tokens[3].c_str());
    }
    if (tok == "Wall" || tok == "Gravity") return fail("Wall and Gravity data are not supported");
    if (tok != "Bead") return fail("Error reading Bead");
    std::vector<std::vector<double>> cons, diss;
    while (tok == "Bead") {
        std::string name;
        double radius;
        if (!t.next(name)) return fail("Error reading bead name");
        if (!t.value(radius) || radius < 0.0) return fail("Error reading bead interaction radius");
        const size_t k = d.bead_names.size() + 1;
        std::vector<double> c(k), g(k);
        for (size_t i = 0; i < k; ++i)
            if (!t.value(c[i])) return fail("Error reading bead conservative value");
        for (size_t i = 0; i < k; ++i)
            if (!t.value(g[i])) return fail("Error reading bead dissipation value");
        if (d.type_of(name) >= 0) return fail("Error adding bead type to CInputData: " + name);
        d.bead_names.push_back(name);
        d.bead_radius.push_back(radius);
        cons.push_back(c);
        diss.push_back(g);
        if (!t.next(tok)) return fail("Error reading Bead token");
    }
    // symmetric matrices from the lower-triangular rows (CInputData::AddBeadType, :505-512)
    const int nt = d.ntypes();
    d.a.assign(nt * nt, 0.0);
    d.g.assign(nt * nt, 0.0);
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j <= i; ++j) {
            d.a[i * nt + j] = d.a[j * nt + i] = cons[i][j];
            d.g[i * nt + j] = d.g[j * nt + i] = diss[i][j];
        }
    while (tok == "Bond") {
        BondType b;
        if (!t.next(b.head) || !t.next(b.tail)) return fail("Error reading bond bead names");
        if (!t.value(b.k) || b.k < 0.0) return fail("Error reading bond spring constant");
        if (!t.value(b.l0) || b.l0 < 0.0) return fail("Error reading bond unstretched length");
        d.bonds.push_back(b);
        if (!t.next(tok)) return fail("Error reading Bond token");
    }
    while (tok == "BondPair") {
        BondPairType b;
        if (!t.next(b.n1) || !t.next(b.n2) || !t.next(b.n3)) return fail("Error reading bond pair bead names");
        if (!t.value(b.k) || b.k < 0.0) return fail("Error reading bond pair bending modulus");
        if (!t.value(b.phi0_deg) || b.phi0_deg < 0.0 || b.phi0_deg > 180.0)
            return fail("Error reading bond pair relaxed angle");
        d.bondpairs.push_back(b);
        if (!t.next(tok)) return fail("Error reading BondPair token");
    }
    if (tok != "Polymer") return fail("Error reading Polymer token");
    double fsum = 0.0;
    bool first = true;
    while (tok == "Polymer") {
        PolymerType p;
        if (!t.next(p.name)) return fail("Error reading polymer name");
        if (!t.value(p.fraction) || p.fraction < 0.0 || p.fraction > 1.0 || (first && p.fraction == 0.0))
            return fail("Error reading polymer number fraction");
        fsum += p.fraction;
        if (fsum > 1.00000001) return fail("Error polymer fractions sum > 1.00000001");
        if (!quoted(t, p.shape)) return fail("Error reading polymer shape");
        first = false;
        if (!t.next(tok)) return fail("Error reading Box token");
        if (tok == "Head") {
            if (!t.next(p.head_name) || !t.next(tok) || tok != "Tail" || !t.next(p.tail_name) ||
                p.tail_name == p.head_name)
                return fail("Error reading polymer Head/Tail");
            if (!t.next(tok)) return fail("Error reading Box token");
        }
        d.polymers.push_back(p);
    }
    if (tok != "Box") return fail("Error reading Box token");
    if (!(t.value(d.nx) && t.value(d.ny) && t.value(d.nz) && t.value(d.wx) && t.value(d.wy) && t.value(d.wz)) ||
        d.nx < 1 || d.ny < 1 || d.nz < 1 || d.wx < 1.0 || d.wy < 1.0 || d.wz < 1.0)
        return fail("Error reading simulation box size");
    auto scalar = [&](const char* name, auto& v) {
        return t.next(tok) && tok == name && t.value(v);
    };
    if (!scalar("Density", d.density) || d.density <= 0.0) return fail("Error reading bead density");
    if (!scalar("Temp", d.kT) || d.kT < 0.0) return fail("Error reading temperature");
    if (!scalar("RNGSeed", d.seed)) return fail("Error reading RNG seed");
    if (!scalar("Lambda", d.lambda) || d.lambda < 0.0 || d.lambda > 1.0) return fail("Error reading lambda");
    if (!scalar("Step", d.dt) || d.dt < 0.000001) return fail("Error reading step size");
    if (!scalar("Time", d.total_time) || d.total_time < 1) return fail("Error reading total time");
    if (!scalar("SamplePeriod", d.sample_period) || d.sample_period < 1 || d.sample_period > d.total_time)
        return fail("Error reading sampling period");
    if (!scalar("AnalysisPeriod", d.analysis_period) || d.analysis_period < 1 ||
        d.analysis_period % d.sample_period != 0 || d.analysis_period > d.total_time)
        return fail("Error reading analysis period");
    if (!scalar("DensityPeriod", d.density_period) || d.density_period < 1 ||
        d.density_period % d.sample_period != 0 || d.density_period > d.total_time)
        return fail("Error reading density period");
    if (!scalar("DisplayPeriod", d.display_period) || d.display_period < 1 || d.display_period > d.total_time)
        return fail("Error reading display period");
    if (!scalar("RestartPeriod", d.restart_period) || d.restart_period < 1 || d.restart_period > d.total_time)
        return fail("Error reading restart period");
    if (!t.next(tok) || tok != "Grid" || !(t.value(d.grid[0]) && t.value(d.grid[1]) && t.value(d.grid[2])) ||
        d.grid[0] < 1 || d.grid[1] < 1 || d.grid[2] < 1)
        return fail("Error reading analysis grid size");
    // trailing blocks
    static const std::set<std::string> blocks = {"Process", "Analysis", "Event", "Command"};
    bool have = (bool)t.next(tok);
    int64_t last_time = 1;
    while (have) {
        if (!blocks.count(tok)) return fail("Error unknown Analysis, Event or Command token: " + tok);
        if (tok == "Command") {
            Command c;
            if (!t.next(c.name) || !t.value(c.time) || c.time < last_time)
                return fail("Error reading Command name/time: " + c.name);
            last_time = c.time;
            have = (bool)t.next(tok);
            while (have && !blocks.count(tok)) {
                c.args.push_back(tok);
                have = (bool)t.next(tok);
            }
            d.commands.push_back(c);
        } else {
            have = (bool)t.next(tok);
            while (have && !blocks.count(tok)) have = (bool)t.next(tok);
        }
    }
    return true;
}

bool read_dmpci_file(const std::string& path, InputData& d, std::string& err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return read_dmpci(ss.str(), d, err);
}

bool build_system(const InputData& d, System& s, std::string& err)
{
    s = System();
    // bond and bond-pair type names: both orders (CInputData::AddBondType /
    // AddBondPairType, src/InputData.cpp:556-674)
    std::map<std::string, int> bond_names, pair_names;
    for (size_t i = 0; i < d.bonds.size(); ++i) {
        const BondType& b = d.bonds[i];
        if (d.type_of(b.head) < 0 || d.type_of(b.tail) < 0 || bond_names.count(b.head + b.tail)) {
            err = "Error adding bond type to CInputData";
            return false;
        }
        bond_names[b.head + b.tail] = (int)i;
        if (b.head != b.tail) bond_names[b.tail + b.head] = (int)i;
    }
    for (size_t i = 0; i < d.bondpairs.size(); ++i) {
        const BondPairType& b = d.bondpairs[i];
        if (d.type_of(b.n1) < 0 || d.type_of(b.n2) < 0 || d.type_of(b.n3) < 0 || !bond_names.count(b.n1 + b.n2) ||
            !bond_names.count(b.n2 + b.n3) || pair_names.count(b.n1 + b.n2 + b.n3)) {
            err = "Error adding bondpair type to CInputData";
            return false;
        }
        pair_names[b.n1 + b.n2 + b.n3] = (int)i;
        if (b.n1 != b.n3) pair_names[b.n3 + b.n2 + b.n1] = (int)i;
    }
    std::vector<PolymerType> types = d.polymers;
    for (PolymerType& p : types) {
        Node root;
        size_t pos = p.shape.find('(');
        if (pos == std::string::npos || !parse_list(p.shape, pos, root, err)) {
            if (err.empty()) err = "invalid polymer shape";
            return false;
        }
        if (root.kids.empty() || !root.kids[0].leaf || !is_name(root.kids[0].name)) {
            err = "polymer shape must start with a bead name: " + p.shape;
            return false;
        }
        Expander ex{d, bond_names, p, {}, ""};
        std::string cur_name;
        int cur = -1;
        if (!ex.seq(root.kids, 0, cur_name, cur)) {
            err = ex.err;
            return false;
        }
        // bond pairs: bond A then every bond B whose head is A's tail; the pair stores
        // (first = B, second = A) (src/InputData.cpp:1114-1164)
        for (size_t A = 0; A < p.bond_head.size(); ++A) {
            const std::string n1 = d.bead_names[p.bead_type[p.bond_head[A]]];
            const std::string n2 = d.bead_names[p.bead_type[p.bond_tail[A]]];
            for (size_t Bi = 0; Bi < p.bond_head.size(); ++Bi) {
                if (p.bond_head[Bi] != p.bond_tail[A]) continue;
                const std::string n3 = d.bead_names[p.bead_type[p.bond_tail[Bi]]];
                auto it = pair_names.find(n1 + n2 + n3);
                if (it == pair_names.end()) continue;
                p.pair_first.push_back((int)Bi);
                p.pair_second.push_back((int)A);
                p.pair_type.push_back(it->second);
            }
        }
    }
    s.types = types;
    // CInitialState::CreatePolymers (src/InitialState.cpp:451-521)
    const double total = d.density * (d.nx * d.wx) * (d.ny * d.wy) * (d.nz * d.wz);
    double nalphasum = 0.0;
    for (const PolymerType& p : types) nalphasum += (double)p.bead_type.size() * p.fraction;
    int32_t bead_id = 0;
    for (size_t pt = 0; pt < types.size(); ++pt) {
        const PolymerType& p = types[pt];
        const int64_t count = (int64_t)(p.fraction * total / nalphasum);
        s.polymers_per_type.push_back(count);
        for (int64_t q = 0; q < count; ++q) {
            const int32_t first_bead = (int32_t)s.id.size();
            const int32_t first_bond = (int32_t)s.bond_head.size();
            ++s.npolymers;
            for (int b : p.bead_type) {
                s.id.push_back(++bead_id);
                s.type.push_back(b);
                s.poly_id.push_back((int32_t)s.npolymers);
                s.poly_type.push_back((int32_t)pt);
            }
            for (size_t b = 0; b < p.bond_head.size(); ++b) {
                s.bond_head.push_back(first_bead + p.bond_head[b]);
                s.bond_tail.push_back(first_bead + p.bond_tail[b]);
                s.bond_k.push_back(d.bonds[p.bond_type[b]].k);
                s.bond_l0.push_back(d.bonds[p.bond_type[b]].l0);
                s.bond_type.push_back(p.bond_type[b]);
            }
            for (size_t b = 0; b < p.pair_first.size(); ++b) {
                s.pair_first.push_back(first_bond + p.pair_first[b]);
                s.pair_second.push_back(first_bond + p.pair_second[b]);
                const BondPairType& bp = d.bondpairs[p.pair_type[b]];
                s.pair_k.push_back(bp.k);
                s.pair_type.push_back(p.pair_type[b]);
                // CBondPairType: degrees to radians with the reference's pi
                s.pair_phi0.push_back(bp.phi0_deg * 3.141592653589793238 / 180.0);
            }
        }
    }
    return true;
}

}  // namespace host
}  // namespace dpd
