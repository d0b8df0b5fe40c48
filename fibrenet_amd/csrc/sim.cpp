// sim.cpp - the simulation host: SimulationManager's driver over the HIP
// engine.
//
// fn_sim uploads a loaded input through the fibrenet::HipManager facade and
// binds every kernel argument as SimulationManager::ReadKernelData does; it
// runs RunSimulationLoop (per-kernel launches with a finish every step, or the
// fused fast path), MinimizeSystemEnergy, and Run: the output_freq chunks with
// their XYZ frames and restart stages.  The writers keep the reference's text
// formats: XYZ frames in fixed notation with 5 decimals, restart frames at the
// stream's default 6 significant digits.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "fibrenet.h"
#include "fibrenet_manager.hpp"
#include "model.h"

struct fn_sim {
    fn_model m;
    fibrenet::HipManager mgr;
    std::string err;
    fn_log_fn log_fn = nullptr;
    void *log_user = nullptr;
    std::string pending;   // messages of fn_sim_create, reported if it fails
    bool creating = true;
    int id_bead = -1, id_info = -1, id_spring = -1, id_angle = -1, id_pair = -1;
    int id_head = -1, id_tail = -1, id_cellnh = -1, id_nh_head = -1, id_nh_tail = -1, id_coeff = -1;
    int beads_num = 0, springs_num = 0, angles_num = 0;
    // DumpData's state, which the reference keeps in function statics
    // (SimulationManager.cpp:1940-1946): frames written, the next restart
    // stage, the exported beads and the open frame files.
    int dump_counter = 1;
    float restart_stage = 0.25f;
    std::vector<int> export_index;
    std::vector<std::string> export_name;
    std::ofstream xyz[3];

    void log(const std::string &s)
    {
        if (creating) pending += s + "\n";
        if (log_fn) log_fn(s.c_str(), log_user);
        else if (!creating) fprintf(stderr, "%s\n", s.c_str());
    }
    bool fail(const std::string &what)
    {
        err = what;
        log("\t" + what);
        return false;
    }
};

namespace {

const char *kFieldSuffix[3] = {"_coord.xyz", "_vel.xyz", "_force.xyz"};

const fn_vector3 &field_of(const fn_bead &b, int field)
{
    return field == 0 ? b.position : field == 1 ? b.velocity : b.force;
}

// The beads DumpData exports (InitializeDataDump, SimulationManager.cpp:1108-1145):
// those whose type is on the export list, in bead order, with their type names.
void export_list(const fn_model &m, std::vector<int> &idx, std::vector<std::string> &names)
{
    idx.clear();
    names.clear();
    for (int i = 0; i < (int)m.info.size(); ++i)
        for (int t : m.export_types)
            if (m.info[i].type == t) {
                idx.push_back(i);
                names.push_back(m.type_name(t));
                break;
            }
}

// One frame of DumpData (SimulationManager.cpp:2002-2048) for one field; the
// stream is in fixed notation with 5 decimals.
void xyz_frame(std::ostream &o, const fn_model &m, const std::vector<int> &idx, const std::vector<std::string> &names,
               const fn_bead *b, int progress, int field)
{
    o << idx.size() << std::endl;
    o << "FRAME" << "\t" << progress << "\t" << m.s.dt << "\t" << m.s.box_size.x << "\t" << m.s.box_size.y << "\t"
      << m.s.box_size.z << std::endl;
    for (size_t k = 0; k < idx.size(); ++k) {
        const fn_vector3 &v = field_of(b[idx[k]], field);
        o << names[k] << "\t" << v.x << "\t" << v.y << "\t" << v.z << std::endl;
    }
}

// DumpRestartData's frame (SimulationManager.cpp:1906-1924) at the stream's
// default formatting.
void restart_frame(std::ostream &o, const fn_model &m, const fn_bead *b, int progress)
{
    const int n = (int)m.beads.size();
    o << n << std::endl;
    o << "FRAME" << "\t" << progress << "\t" << m.s.dt << "\t" << m.s.box_size.x << "\t" << m.s.box_size.y << "\t"
      << m.s.box_size.z << std::endl;
    for (int i = 0; i < n; ++i) {
        const fn_bead &r = b[i];
        o << "\t" << m.type_name(m.info[i].type);
        o << "\t" << r.position.x << "\t" << r.position.y << "\t" << r.position.z;
        o << "\t" << r.image_position.x << "\t" << r.image_position.y << "\t" << r.image_position.z;
        o << "\t" << r.velocity.x << "\t" << r.velocity.y << "\t" << r.velocity.z;
        o << "\t" << r.force.x << "\t" << r.force.y << "\t" << r.force.z << std::endl;
    }
}

bool write_restart_file(const fn_model &m, const fn_bead *b, int progress, const std::string &path)
{
    std::ofstream f(path, std::ios::out | std::ios::trunc);
    if (!f.is_open() || !f.good()) return false;
    restart_frame(f, m, b, progress);
    return f.good();
}

// SimulationManager::ReadKernelData (SimulationManager.cpp:1180-1307).
bool read_kernel_data(fn_sim &S)
{
    auto &M = S.mgr;
    std::vector<std::string> names = {"AdvancePosition", "ResetGrid", "ResetContactList", "FillGrid",
                                      "BuildContactList", "ResolveContacts", "AdvanceVelocity",
                                      "ResetBeadEnergy", "ResolveSprings", "ResolveAngles"};
    if (!M.RegisterKernel(names)) return false;
    auto bead = M.GetBuffer(S.id_bead), angle = M.GetBuffer(S.id_angle), spring = M.GetBuffer(S.id_spring);
    auto info = M.GetBuffer(S.id_info), head = M.GetBuffer(S.id_head), tail = M.GetBuffer(S.id_tail);
    auto cellnh = M.GetBuffer(S.id_cellnh), nh_head = M.GetBuffer(S.id_nh_head), nh_tail = M.GetBuffer(S.id_nh_tail);
    auto coeff = M.GetBuffer(S.id_coeff), table = M.GetBuffer(S.id_pair);
    if (!bead || !head || !tail || !cellnh || !nh_head || !nh_tail || !coeff || !table) return false;
    const fn_settings &settings = S.m.s;
    int beads_num = M.GetBufferElements(S.id_bead);
    int grid_size = M.GetBufferElements(S.id_head);
    int list_size = M.GetBufferElements(S.id_nh_head);
    bool ok = true;
    ok &= M.SetKernelArg("AdvancePosition", 0, *bead);
    ok &= M.SetKernelArg("AdvancePosition", 1, *info);
    ok &= M.SetKernelArg("AdvancePosition", 2, settings);
    ok &= M.SetKernelArg("AdvancePosition", 3, beads_num);
    M.SetKernelGlobalWorkRange("AdvancePosition", M.GetBufferLength(S.id_bead));
    if (spring) {
        int springs_num = M.GetBufferElements(S.id_spring);
        ok &= M.SetKernelArg("ResolveSprings", 0, *bead);
        ok &= M.SetKernelArg("ResolveSprings", 1, *spring);
        ok &= M.SetKernelArg("ResolveSprings", 2, springs_num);
        M.SetKernelGlobalWorkRange("ResolveSprings", M.GetBufferLength(S.id_spring));
    }
    if (angle) {
        ok &= M.SetKernelArg("ResolveAngles", 0, *angle);
        ok &= M.SetKernelArg("ResolveAngles", 1, *spring);
        ok &= M.SetKernelArg("ResolveAngles", 2, *bead);
        ok &= M.SetKernelArg("ResolveAngles", 3, S.angles_num);
        M.SetKernelGlobalWorkRange("ResolveAngles", M.GetBufferLength(S.id_angle));
    }
    ok &= M.SetKernelArg("ResetGrid", 0, *head);
    ok &= M.SetKernelArg("ResetGrid", 1, grid_size);
    M.SetKernelGlobalWorkRange("ResetGrid", M.GetBufferLength(S.id_head));
    ok &= M.SetKernelArg("ResetContactList", 0, *nh_head);
    ok &= M.SetKernelArg("ResetContactList", 1, list_size);
    M.SetKernelGlobalWorkRange("ResetContactList", M.GetBufferLength(S.id_nh_head));
    ok &= M.SetKernelArg("FillGrid", 0, *bead);
    ok &= M.SetKernelArg("FillGrid", 1, *head);
    ok &= M.SetKernelArg("FillGrid", 2, *tail);
    ok &= M.SetKernelArg("FillGrid", 3, settings);
    ok &= M.SetKernelArg("FillGrid", 4, beads_num);
    M.SetKernelGlobalWorkRange("FillGrid", M.GetBufferLength(S.id_bead));
    ok &= M.SetKernelArg("BuildContactList", 0, *head);
    ok &= M.SetKernelArg("BuildContactList", 1, *tail);
    ok &= M.SetKernelArg("BuildContactList", 2, *cellnh);
    ok &= M.SetKernelArg("BuildContactList", 3, *nh_head);
    ok &= M.SetKernelArg("BuildContactList", 4, *nh_tail);
    ok &= M.SetKernelArg("BuildContactList", 5, *bead);
    ok &= M.SetKernelArg("BuildContactList", 6, *info);
    ok &= M.SetKernelArg("BuildContactList", 7, *coeff);
    ok &= M.SetKernelArg("BuildContactList", 8, *table);
    ok &= M.SetKernelArg("BuildContactList", 9, settings);
    ok &= M.SetKernelArg("BuildContactList", 10, grid_size);
    M.SetKernelGlobalWorkRange("BuildContactList", M.GetBufferLength(S.id_head));
    ok &= M.SetKernelArg("ResolveContacts", 0, *nh_head);
    ok &= M.SetKernelArg("ResolveContacts", 1, *nh_tail);
    ok &= M.SetKernelArg("ResolveContacts", 2, *bead);
    ok &= M.SetKernelArg("ResolveContacts", 3, *info);
    ok &= M.SetKernelArg("ResolveContacts", 4, *coeff);
    ok &= M.SetKernelArg("ResolveContacts", 5, *table);
    ok &= M.SetKernelArg("ResolveContacts", 6, settings);
    ok &= M.SetKernelArg("ResolveContacts", 7, list_size);
    M.SetKernelGlobalWorkRange("ResolveContacts", M.GetBufferLength(S.id_nh_head));
    ok &= M.SetKernelArg("AdvanceVelocity", 0, *bead);
    ok &= M.SetKernelArg("AdvanceVelocity", 1, *info);
    ok &= M.SetKernelArg("AdvanceVelocity", 2, beads_num);
    int padded = M.GetBufferLength(S.id_bead);
    ok &= M.SetKernelArg("AdvanceVelocity", 3, padded);
    M.SetKernelGlobalWorkRange("AdvanceVelocity", M.GetBufferLength(S.id_bead));
    ok &= M.SetKernelArg("ResetBeadEnergy", 0, *bead);
    ok &= M.SetKernelArg("ResetBeadEnergy", 1, beads_num);
    M.SetKernelGlobalWorkRange("ResetBeadEnergy", M.GetBufferLength(S.id_bead));
    return ok;
}

// Buffer creation in the order and with the sizes of the reference's loaders.
bool upload(fn_sim &S)
{
    auto &M = S.mgr;
    fn_model &m = S.m;
    const int C = m.s.cells_x * m.s.cells_y * m.s.cells_z;
    S.id_head = M.CreateBuffer<int>(C, CL_MEM_READ_WRITE);                       // :2304
    S.id_tail = M.CreateBuffer<fn_grid_cell>(C, CL_MEM_READ_WRITE);              // :2312
    std::vector<fn_grid_cell_nh> nh = m.cellnh;
    S.id_cellnh = M.CreateBuffer<fn_grid_cell_nh>(nh, CL_MEM_READ_ONLY);         // :2319
    std::vector<fn_pair> pairs = m.pairs;
    S.id_pair = M.CreateBuffer<fn_pair>(pairs, CL_MEM_READ_ONLY);                // :963
    std::vector<fn_bead> beads = m.beads;
    std::vector<fn_bead_info> info = m.info;
    S.beads_num = (int)beads.size();
    S.id_bead = M.CreateBuffer<fn_bead>(beads, CL_MEM_READ_WRITE);               // :1461
    S.id_info = M.CreateBuffer<fn_bead_info>(info, CL_MEM_READ_ONLY);            // :1467
    S.id_nh_head = M.CreateBuffer<int>(S.beads_num, CL_MEM_READ_WRITE);          // :1473
    S.id_nh_tail = M.CreateBuffer<fn_particle_nh>(S.beads_num, CL_MEM_READ_WRITE);  // :1478
    S.id_coeff = M.CreateBuffer<fn_particle_nh>(S.beads_num, CL_MEM_READ_WRITE);    // :1483
    if (!m.springs.empty()) {
        std::vector<fn_spring> sp = m.springs;
        S.springs_num = (int)sp.size();
        S.id_spring = M.CreateBuffer<fn_spring>(sp, CL_MEM_READ_WRITE);          // :1572
    }
    if (!m.angles.empty()) {
        std::vector<fn_angle> an = m.angles;
        S.angles_num = (int)an.size();
        S.id_angle = M.CreateBuffer<fn_angle>(an, CL_MEM_READ_ONLY);             // :1650
    }
    return S.id_head >= 0 && S.id_tail >= 0 && S.id_cellnh >= 0 && S.id_pair >= 0 && S.id_bead >= 0 &&
           S.id_info >= 0 && S.id_nh_head >= 0 && S.id_nh_tail >= 0 && S.id_coeff >= 0 &&
           (m.springs.empty() || S.id_spring >= 0) && (m.angles.empty() || S.id_angle >= 0);
}

// SimulationManager::RunSimulationLoop (SimulationManager.cpp:151-181).
bool run_loop(fn_sim &S, int loop_size, bool fast)
{
    if (loop_size <= 0) {
        S.err = "RunSimulationLoop: loop size should be a positive number";
        return false;
    }
    auto &M = S.mgr;
    if (fast) return M.RunSteps(loop_size, S.m.grid_update_freq, S.springs_num > 0, S.angles_num > 0);
    int cnt = 0;
    for (int step = 0; step < loop_size; ++step) {
        if (cnt == S.m.grid_update_freq) cnt = 0;
        if (!M.ExecuteKernel("AdvancePosition")) return false;
        if (S.springs_num > 0 && !M.ExecuteKernel("ResolveSprings")) return false;
        if (S.angles_num > 0 && !M.ExecuteKernel("ResolveAngles")) return false;
        if (cnt == 0) {
            if (!M.ExecuteKernel("ResetGrid")) return false;
            if (!M.ExecuteKernel("ResetContactList")) return false;
            if (!M.ExecuteKernel("FillGrid")) return false;
            if (!M.ExecuteKernel("BuildContactList")) return false;
        }
        if (!M.ExecuteKernel("ResolveContacts")) return false;
        if (!M.ExecuteKernel("AdvanceVelocity")) return false;
        if (!M.RunKernels()) return false;
        ++cnt;
    }
    return true;
}



// The bead records as the device holds them now.
bool read_beads(fn_sim &S, std::vector<fn_bead> &out)
{
    out.assign(S.mgr.GetBufferLength(S.id_bead), fn_bead{});
    if (out.empty() || !S.mgr.ReadBuffer(S.id_bead, out.data())) return S.fail("unable to read the bead buffer");
    return true;
}

// SimulationManager::DumpRestartData (SimulationManager.cpp:1855-1928).
bool dump_restart(fn_sim &S, int progress, const std::string &path)
{
    std::vector<fn_bead> b;
    if (!read_beads(S, b)) return false;
    if (!write_restart_file(S.m, b.data(), progress, path))
        return S.fail("DumpRestartData: ERROR - unable to write the restart file " + path);
    S.log("\tDumpRestartData: restart data dump successful.");
    return true;
}

// SimulationManager::DumpData (SimulationManager.cpp:1937-2061): one frame per
// dumped field, appended to <output_file>_coord / _vel / _force.xyz, and the
// restart file whenever progress / run passes the next quarter.
bool dump_data(fn_sim &S, int progress)
{
    const fn_model &m = S.m;
    const bool on[3] = {m.export_position, m.export_velocity, m.export_force};
    if (S.dump_counter == 1) {
        export_list(m, S.export_index, S.export_name);
        if (S.export_index.empty())
            return S.fail("InitializeDataDump: ERROR - declared export types not found in model - there is no data "
                          "to export.");
        for (int f = 0; f < 3; ++f)
            if (on[f]) {
                S.xyz[f].open(m.output_file + kFieldSuffix[f], std::ios_base::app);
                if (!S.xyz[f].good()) S.log("\tDumpData: ERROR - unable to initialize " + m.output_file + kFieldSuffix[f]);
                S.xyz[f] << std::fixed << std::setprecision(5);
            }
    }
    std::vector<fn_bead> b;
    if (!read_beads(S, b)) return false;
    for (int f = 0; f < 3; ++f) {
        if (!on[f]) continue;
        if (!S.xyz[f].is_open() || !S.xyz[f].good())
            return S.fail(std::string("DumpData: ERROR - unable to write data to ") + m.output_file + kFieldSuffix[f]);
        xyz_frame(S.xyz[f], m, S.export_index, S.export_name, b.data(), progress, f);
        S.xyz[f].flush();
    }
    if ((static_cast<float>(progress) / static_cast<float>(m.run)) > S.restart_stage) {
        S.restart_stage += 0.25f;
        if (!write_restart_file(m, b.data(), progress, m.output_file + "_restart.res"))
            return S.fail("DumpData: ERROR - unable to dump restart data at step " + std::to_string(progress));
        S.log("\tDumpRestartData: restart data dump successful.");
    }
    S.dump_counter++;
    return true;
}

// SimulationManager::MinimizeSystemEnergy (SimulationManager.cpp:111-143).
bool minimize(fn_sim &S, bool fast)
{
    fn_model &m = S.m;
    S.log("\tMinimizeSystemEnergy: initializing system energy minimalization.");
    if (m.min_steps < 0) return S.fail("MinimizeSystemEnergy: ERROR - loop size should be a positive number.");
    if (m.min_reset && !S.mgr.ExecuteKernel("ResetBeadEnergy")) return S.fail("MinimizeSystemEnergy: reset failed");
    if (m.min_steps == 0) {
        S.log("\tMinimizeSystemEnergy: number of minimalization steps set to zero, skipping minimalization.");
        return true;
    }
    if (m.min_sub <= 0) return S.fail("MinimizeSystemEnergy: minimize_substeps is zero (the reference divides by it)");
    const int loops = m.min_steps / m.min_sub;
    for (int p = 0; p < loops; ++p) {
        if (!run_loop(S, m.min_sub, fast)) return false;
        S.log("\tMinimizeSystemEnergy: minimalization progress - finished " + std::to_string(p * m.min_sub) +
              " out of " + std::to_string(m.min_sub * loops) + " minimalization steps");
        if (m.min_reset && !S.mgr.ExecuteKernel("ResetBeadEnergy"))
            return S.fail("MinimizeSystemEnergy: reset failed");
    }
    if (!S.mgr.RunKernels()) return S.fail("MinimizeSystemEnergy: the device failed");
    S.log("\tMinimizeSystemEnergy: system energy minimalization finished.");
    return true;
}

template <typename X>
std::string str(const X &v)
{
    std::ostringstream o;
    o << v;
    return o.str();
}

// SimulationManager::Run (SimulationManager.cpp:54-103).
bool run(fn_sim &S, bool fast)
{
    const fn_model &m = S.m;
    const fn_settings &s = m.s;
    S.log("");
    S.log("%-------  Running Simulation  -------%");
    S.log("\tBox size:\t" + str(s.box_size.x) + " " + str(s.box_size.y) + " " + str(s.box_size.z));
    S.log("\tGrid cell size:\t" + str(s.grid_step.x) + " x " + str(s.grid_step.y) + " x " + str(s.grid_step.z));
    S.log("\tGrid cells:\t" + str(s.cells_x) + " x " + str(s.cells_y) + " x " + str(s.cells_z));
    S.log("\tSimulation steps:\t" + str(m.run));
    S.log("\tData dump frequency:\t" + str(m.dump));
    S.log("\tTime step:\t" + str(s.dt));
    const auto t0 = std::chrono::steady_clock::now();
    if (m.min_steps > 0 && !minimize(S, fast)) return false;
    int progress = 0;
    if (!dump_data(S, progress)) return false;
    while (progress < m.run) {
        const int chunk = (m.run - progress) < m.dump ? m.run - progress : m.dump;
        if (!run_loop(S, chunk, fast)) return S.fail(S.err.empty() ? "RunSimulationLoop failed" : S.err);
        progress += chunk;
        if (!dump_data(S, progress)) return false;
        S.log("\tSimulation progress: finished " + str(progress) + " out of " + str(m.run) + " simulation steps");
    }
    const double minutes =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 60.0;
    if (!dump_restart(S, progress, m.output_file + "_restart.res"))
        return S.fail("Error: unable to dump final restart data frame.");
    S.log("\tSimulation finished after " + str(minutes) + " minutes");
    return true;
}

}  // namespace

extern "C" {

int fn_sim_create(int device, int precision, const char *input_path, fn_sim **out)
{
    if (!input_path || !out) return FN_ERR_ARG;
    *out = nullptr;
    fn_sim *S = new fn_sim();
    try {
        fnload::load_all(S->m, input_path);
    } catch (const std::exception &e) {
        fnload::set_error(e.what());
        delete S;
        return FN_ERR_LOAD;
    }
    if (!S->mgr.Initialize([S](const std::string &s) { S->log(s); }, precision, device)) {
        fnload::set_error("unable to open the HIP device");
        delete S;
        return FN_ERR_HIP;
    }
    if (!upload(*S) || !read_kernel_data(*S)) {
        fnload::set_error("upload/bind failed: " + S->pending);
        delete S;
        return FN_ERR_ARG;
    }
    S->creating = false;
    S->pending.clear();
    fnload::set_error("");
    *out = S;
    return FN_OK;
}

int fn_sim_destroy(fn_sim *sim)
{
    delete sim;
    return FN_OK;
}

fn_engine *fn_sim_engine(fn_sim *sim) { return sim ? sim->mgr.engine() : nullptr; }

int fn_sim_buffer_id(fn_sim *sim, int kind)
{
    if (!sim) return FN_ERR_ARG;
    switch (kind) {
    case FN_BUF_BEAD: return sim->id_bead;
    case FN_BUF_BEAD_INFO: return sim->id_info;
    case FN_BUF_SPRING: return sim->id_spring;
    case FN_BUF_ANGLE: return sim->id_angle;
    case FN_BUF_PAIR: return sim->id_pair;
    default: return FN_ERR_ARG;
    }
}

int fn_sim_run_loop(fn_sim *sim, int loop_size, int fast)
{
    if (!sim) return FN_ERR_ARG;
    sim->err.clear();
    return run_loop(*sim, loop_size, fast != 0) ? FN_OK : FN_ERR_ARG;
}

int fn_sim_minimize(fn_sim *sim, int fast)
{
    if (!sim) return FN_ERR_ARG;
    sim->err.clear();
    return minimize(*sim, fast != 0) ? FN_OK : FN_ERR_ARG;
}

int fn_sim_run(fn_sim *sim, int fast)
{
    if (!sim) return FN_ERR_ARG;
    sim->err.clear();
    return run(*sim, fast != 0) ? FN_OK : FN_ERR_ARG;
}

int fn_sim_dump(fn_sim *sim, int progress)
{
    if (!sim) return FN_ERR_ARG;
    sim->err.clear();
    return dump_data(*sim, progress) ? FN_OK : FN_ERR_ARG;
}

int fn_sim_write_restart(fn_sim *sim, const char *path, int progress)
{
    if (!sim) return FN_ERR_ARG;
    sim->err.clear();
    return dump_restart(*sim, progress, path ? std::string(path) : sim->m.output_file + "_restart.res") ? FN_OK
                                                                                                      : FN_ERR_ARG;
}

int fn_sim_set_log_callback(fn_sim *sim, fn_log_fn fn, void *user)
{
    if (!sim) return FN_ERR_ARG;
    sim->log_fn = fn;
    sim->log_user = user;
    return FN_OK;
}

const char *fn_sim_error(fn_sim *sim)
{
    if (!sim) return "";
    if (!sim->err.empty()) return sim->err.c_str();
    fn_engine *e = sim->mgr.engine();
    return e ? fn_last_error(e) : "";
}

const char *fn_sim_output_file(fn_sim *sim) { return sim ? sim->m.output_file.c_str() : ""; }

int fn_model_write_restart(const fn_model *m, const fn_bead *beads, int progress, const char *path)
{
    if (!m || !beads || !path) return FN_ERR_ARG;
    return write_restart_file(*m, beads, progress, path) ? FN_OK : FN_ERR_ARG;
}

int fn_model_write_xyz(const fn_model *m, const fn_bead *beads, int progress, int field, const char *path)
{
    if (!m || !beads || !path || field < FN_OUT_POSITION || field > FN_OUT_FORCE) return FN_ERR_ARG;
    std::vector<int> idx;
    std::vector<std::string> names;
    export_list(*m, idx, names);
    std::ofstream f(path, std::ios_base::app);
    if (!f.good()) return FN_ERR_ARG;
    f << std::fixed << std::setprecision(5);
    xyz_frame(f, *m, idx, names, beads, progress, field);
    return f.good() ? FN_OK : FN_ERR_ARG;
}

}  // extern "C"
