// dmpci.hpp -- host model of an Osprey-DPD control file (dmpci.<runId>) and the
// polymer topology it defines.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace dpd {
namespace host {

struct BondType {
    std::string head, tail;
    double k = 0, l0 = 0;
};

struct BondPairType {
    std::string n1, n2, n3;
    double k = 0, phi0_deg = 0;
};

// One polymer template, expanded from its shape string exactly as
// CInputData::AddPolymerType / ParseElement / AddBeadtoPolymer do
// (src/InputData.cpp:1028-1188, :1469-1513, :1711-1790).
struct PolymerType {
    std::string name, shape, head_name, tail_name;
    double fraction = 0;
    std::vector<int> bead_type;               // per bead in creation order
    std::vector<int> bond_head, bond_tail;    // bead indices (head = earlier bead)
    std::vector<int> bond_type;
    std::vector<int> pair_first, pair_second; // bond indices (first = later bond)
    std::vector<int> pair_type;
};

struct Command {
    std::string name;
    int64_t time = 0;
    std::vector<std::string> args;
};

struct InputData {
    std::string title, date, comment;
    std::string state;                         // random, restart, lamella, ...
    std::vector<std::string> state_tokens;     // data of the State block, verbatim
    std::string restart_run;
    int64_t restart_state = 0;
    std::vector<std::string> bead_names;
    std::vector<double> bead_radius;
    std::vector<double> a, g;                  // ntypes^2 symmetric
    std::vector<BondType> bonds;
    std::vector<BondPairType> bondpairs;
    std::vector<PolymerType> polymers;
    int nx = 0, ny = 0, nz = 0;
    double wx = 1, wy = 1, wz = 1;
    double density = 0, kT = 0, lambda = 0, dt = 0;
    int64_t seed = 0, total_time = 0;
    int64_t sample_period = 0, analysis_period = 0, density_period = 0, display_period = 0,
            restart_period = 0;
    int grid[3] = {1, 1, 1};
    std::vector<Command> commands;
    int ntypes() const { return (int)bead_names.size(); }
    // values following `key` in the State block (e.g. "Normal" -> {"0","0","1"})
    std::vector<std::string> state_values(const std::string& key, size_t count) const;
    int type_of(const std::string& name) const;
};

// Parse a dmpci file (CInputDPDFile::Serialize, src/InputDPDFile.cpp:190-1052).  Returns
// false with a message naming the failing item, as the reference's IOError does.
bool read_dmpci(const std::string& text, InputData& d, std::string& err);
bool read_dmpci_file(const std::string& path, InputData& d, std::string& err);

// Expanded system: polymers created in type order with counts
// floor(fraction * rho V / sum(size * fraction)) (CInitialState::CreatePolymers,
// src/InitialState.cpp:451-521); bead ids 1.. in creation order.
struct System {
    std::vector<int32_t> id, type, poly_id, poly_type;
    std::vector<int32_t> bond_head, bond_tail;
    std::vector<double> bond_k, bond_l0;
    std::vector<int32_t> pair_first, pair_second;
    std::vector<double> pair_k, pair_phi0;   // radians (src/BondPairType.cpp:85)
    std::vector<int32_t> bond_type, pair_type;   // declaration index of each bond / bond pair
    std::vector<int64_t> polymers_per_type;
    std::vector<PolymerType> types;          // expanded polymer templates
    int64_t npolymers = 0;
};

bool build_system(const InputData& d, System& s, std::string& err);

}  // namespace host
}  // namespace dpd
