// model.h - host records of a FibreNet input (fn_model), shared by the
// loaders (loader.cpp) and the simulation host (sim.cpp).
#pragma once
#include <map>
#include <string>
#include <vector>

#include "fibrenet.h"

struct fn_model {
    fn_settings s{};
    std::map<std::string, int> type_id;
    std::vector<fn_pair> pairs;
    std::vector<fn_bead> beads;
    std::vector<fn_bead_info> info;
    std::vector<fn_spring> springs;
    std::vector<fn_angle> angles;
    std::vector<fn_grid_cell_nh> cellnh;
    int grid_update_freq = 8;
    int run = -1, dump = -1, min_steps = 0, min_sub = 0;
    bool min_reset = false, restart = false;
    double skin_radius_ratio = 0.5;
    // Output settings of LoadExportInfo (SimulationManager.cpp:979-1091): the
    // output_file name (resolved against the input file's directory), the
    // exported bead types and which of position / velocity / force are dumped.
    std::string output_file;
    std::vector<int> export_types;
    bool export_position = false, export_velocity = false, export_force = false;

    // Name of bead type t (the first name mapped to t, as the reference's
    // std::find_if over its type map finds it).
    const std::string &type_name(int t) const
    {
        static const std::string none = "?";
        for (const auto &kv : type_id)
            if (kv.second == t) return kv.first;
        return none;
    }
};

namespace fnload {
// SimulationManager::ReadInputFile (SimulationManager.cpp:1153-1171): the
// input file, its model and restart files, modifiers; throws on rejection.
void load_all(fn_model &m, const std::string &path);
// Sets the message fn_model_error() returns.
void set_error(const std::string &message);
}  // namespace fnload
