// loader.cpp - FibreNet input/model file loaders.
//
// The loaders keep the reference's file formats and acceptance rules
// (SimulationManager.cpp, cited per function): keywords are found by the first
// token of a line (TextParser.h:21-38), single-valued keys take their first
// occurrence, numbers go through std::stod / std::stoi / std::stof (so "1e3"
// read by stoi is 1 and a rest angle is float-rounded), and the same inputs
// are rejected.  Fields the reference leaves uninitialised are zero here.
//
// Modifiers (modif_bead / modif_spring) are applied to the host records before
// the first upload; the reference reads the device buffers back, edits and
// writes them, which is the same result for an fp64 device.
//
// Paths named in the input file (model, restart, output) are resolved against
// the input file's directory when relative; the reference opens them relative
// to its working directory, which is the same when it is started there.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fibrenet.h"
#include "model.h"

namespace {

struct LoadError : std::runtime_error {
    explicit LoadError(const std::string &m) : std::runtime_error(m) {}
};

thread_local std::string g_model_error;

typedef std::vector<std::string> Lines;
typedef std::vector<std::string> Tokens;

// TextParser::ReadFileContent (TextParser.h:7-19).
bool read_lines(const std::string &path, Lines &out)
{
    std::ifstream f(path);
    if (!f.is_open()) return false;
    out.clear();
    std::string s;
    while (std::getline(f, s)) out.push_back(s);
    return true;
}

Tokens tokens_of(const std::string &line)
{
    std::istringstream iss(line);
    return Tokens(std::istream_iterator<std::string>(iss), std::istream_iterator<std::string>());
}

// TextParser::ScanText (TextParser.h:21-38): 1-based line of the first line at
// or after index `start` whose first token is `key`, or -1.
int scan(const Lines &text, const std::string &key, Tokens &tok, int start)
{
    tok.clear();
    if (start < 0) return -1;
    for (int ii = start; ii < (int)text.size(); ++ii) {
        tok = tokens_of(text[ii]);
        if (!tok.empty() && tok[0] == key) return ii + 1;
        tok.clear();
    }
    return -1;
}

double stod_(const std::string &t)
{
    try { return std::stod(t); } catch (const std::exception &) { throw LoadError("invalid number '" + t + "'"); }
}
int stoi_(const std::string &t)
{
    try { return std::stoi(t); } catch (const std::exception &) { throw LoadError("invalid integer '" + t + "'"); }
}
float stof_(const std::string &t)
{
    try { return std::stof(t); } catch (const std::exception &) { throw LoadError("invalid number '" + t + "'"); }
}

}  // namespace


namespace {

// SimulationManager.cpp:644-673 LoadGlobalForceInfo.
void load_global_force(fn_model &m, const Lines &L)
{
    Tokens t;
    if (scan(L, "gravity", t, 0) > 0) {
        if (t.size() != 4) throw LoadError("gravity: incorrect number of command parameters");
        m.s.gravity.x = stod_(t[1]);
        m.s.gravity.y = stod_(t[2]);
        m.s.gravity.z = stod_(t[3]);
    }
    if (scan(L, "damping", t, 0) > 0) {
        if (t.size() != 2) throw LoadError("damping: incorrect number of command parameters");
        m.s.damping = stod_(t[1]) * -1.0f;
    }
}

// SimulationManager.cpp:681-819 LoadSolverInfo.
void load_solver(fn_model &m, const Lines &L)
{
    Tokens t;
    if (scan(L, "step", t, 0) <= 0) throw LoadError("'step' not specified");
    if (t.size() != 2) throw LoadError("step: incorrect number of command parameters");
    m.s.dt = stod_(t[1]);
    if (!(m.s.dt > 0.0f)) throw LoadError("step: incorrect value");
    m.s.half_dt = m.s.dt * 0.5f;
    m.s.half_dt2 = m.s.dt * m.s.dt * 0.5f;
    if (scan(L, "lambda", t, 0) <= 0) throw LoadError("'lambda' not specified (the reference stops here)");
    if (t.size() != 2) throw LoadError("lambda: incorrect number of command parameters");
    m.s.lambda_dt = stod_(t[1]);
    if (!(m.s.lambda_dt > 0.0f)) throw LoadError("lambda: incorrect value");
    m.s.lambda_dt *= m.s.dt;
    if (scan(L, "run", t, 0) <= 0) throw LoadError("'run' not specified");
    if (t.size() != 2) throw LoadError("run: incorrect number of command parameters");
    m.run = stoi_(t[1]);
    if (m.run < 0) throw LoadError("run: incorrect value");
    if (scan(L, "minimize_steps", t, 0) > 0) {
        if (t.size() != 2) throw LoadError("minimize_steps: incorrect number of command parameters");
        m.min_steps = stoi_(t[1]);
        if (m.min_steps < 0) throw LoadError("minimize_steps: incorrect value");
    }
    if (scan(L, "minimize_substeps", t, 0) > 0) {
        if (t.size() != 2) throw LoadError("minimize_substeps: incorrect number of command parameters");
        m.min_sub = stoi_(t[1]);
        if (m.min_sub < 0 || m.min_sub > m.min_steps) throw LoadError("minimize_substeps: incorrect value");
    } else {
        m.min_sub = m.min_steps > 0 ? m.min_steps : 0;
    }
    if (scan(L, "minimize_reset", t, 0) > 0) {
        if (t.size() != 2) throw LoadError("minimize_reset: incorrect number of command parameters");
        if (t[1] == "yes") m.min_reset = true;
        else if (t[1] == "no") m.min_reset = false;
        else throw LoadError("minimize_reset: incorrect value");
    }
    if (scan(L, "cutoff", t, 0) <= 0) throw LoadError("'cutoff' not specified");
    if (t.size() != 2) throw LoadError("cutoff: incorrect number of command parameters");
    m.s.cutoff = stod_(t[1]);
    m.s.cutoff_sq = m.s.cutoff * m.s.cutoff;
    if (m.s.cutoff <= 0.0f) throw LoadError("cutoff: incorrect value");
}

// SimulationManager.cpp:2070-2301 LoadSimulationBoxInfo (grid geometry and the
// 13-cell half-shell table).
void load_box(fn_model &m, const Lines &L)
{
    Tokens t;
    fn_settings &s = m.s;
    if (scan(L, "box_unit_cell", t, 0) <= 0) throw LoadError("'box_unit_cell' not specified");
    if (t.size() != 2) throw LoadError("box_unit_cell: incorrect number of command parameters");
    s.box_unit_cell = stod_(t[1]);
    if (s.box_unit_cell <= 0.0f) throw LoadError("box_unit_cell: incorrect value");
    if (scan(L, "box_size", t, 0) <= 0) throw LoadError("'box_size' not specified");
    if (t.size() != 4) throw LoadError("box_size: incorrect number of command parameters");
    s.box_size.x = stod_(t[1]);
    s.box_size.y = stod_(t[2]);
    s.box_size.z = stod_(t[3]);
    if (s.box_size.x < 0.0f || s.box_size.y < 0.0f || s.box_size.z < 0.0f)
        throw LoadError("box dimensions must have positive values");
    if ((s.box_size.x / s.box_unit_cell) < 3.0f || (s.box_size.y / s.box_unit_cell) < 3.0f ||
        (s.box_size.z / s.box_unit_cell) < 3.0f)
        throw LoadError("box dimensions must be at least three grid cells");
    if (scan(L, "box_type", t, 0) <= 0) throw LoadError("'box_type' not specified");
    if (t.size() != 2) throw LoadError("box_type: incorrect number of command parameters");
    if (t[1] == "pbc") s.bc_type = 0;
    else if (t[1] == "finite") s.bc_type = 1;
    else throw LoadError("box_type: incorrect value");
    if (scan(L, "grid_update_freq", t, 0) > 0) {
        if (t.size() != 2) throw LoadError("grid_update_freq: incorrect number of command parameters");
        m.grid_update_freq = stoi_(t[1]);
        if (m.grid_update_freq <= 0) throw LoadError("grid_update_freq: incorrect value");
    } else {
        m.grid_update_freq = 8;
    }
    if (scan(L, "box_wall_rep", t, 0) > 0) {
        if (t.size() != 2) throw LoadError("box_wall_rep: incorrect number of command parameters");
        s.wall_contact_stiffness = stod_(t[1]);
        if (s.wall_contact_stiffness < 0.0f) throw LoadError("box_wall_rep: incorrect value");
    } else {
        s.bc_type = 1;   // SimulationManager.cpp:2167-2170
    }
    if (s.box_size.x <= 0.0f || s.box_size.y <= 0.0f || s.box_size.z <= 0.0f) throw LoadError("invalid grid size");
    s.cells_x = (int)floor(s.box_size.x / s.box_unit_cell);
    s.cells_y = (int)floor(s.box_size.y / s.box_unit_cell);
    s.cells_z = (int)floor(s.box_size.z / s.box_unit_cell);
    s.cells_xy = s.cells_x * s.cells_y;
    s.grid_step.x = s.box_size.x / (float)s.cells_x;
    s.grid_step.y = s.box_size.y / (float)s.cells_y;
    s.grid_step.z = s.box_size.z / (float)s.cells_z;
    s.half_box_sq = std::min<double>(std::min<double>(s.box_size.x, s.box_size.y), s.box_size.z) * 0.5f;
    s.half_box_sq *= s.half_box_sq;
    const int cx = s.cells_x, cy = s.cells_y, cz = s.cells_z;
    fn_grid_cell_nh none;
    for (int &v : none.n) v = -1;
    m.cellnh.assign((size_t)cx * cy * cz, none);
    auto nb = [&](int x, int y, int z) {
        if (s.bc_type == 0) {
            if (x < 0) x = cx - 1;
            if (x >= cx) x = 0;
            if (y < 0) y = cy - 1;
            if (y >= cy) y = 0;
            if (z < 0) z = cz - 1;
            if (z >= cz) z = 0;
            return x + y * cx + z * cx * cy;
        }
        if (x < 0 || x >= cx || y < 0 || y >= cy || z < 0 || z >= cz) return -1;
        return x + y * cx + z * cx * cy;
    };
    for (int z = 0; z < cz; ++z)
        for (int y = 0; y < cy; ++y)
            for (int x = 0; x < cx; ++x) {
                int *o = m.cellnh[(size_t)x + (size_t)y * cx + (size_t)z * cx * cy].n;
                int k = 0;
                for (int dx = -1; dx < 2; ++dx)
                    for (int dy = -1; dy < 2; ++dy) o[k++] = nb(x + dx, y + dy, z - 1);
                for (int dy = -1; dy < 2; ++dy) o[k++] = nb(x + 1, y + dy, z);
                o[k++] = nb(x, y - 1, z);
            }
}

// SimulationManager.cpp:827-970 LoadTypeAndPairInfo.
void load_types_pairs(fn_model &m, const Lines &L)
{
    Tokens t;
    int ln = 0;
    int T = 0;
    while ((ln = scan(L, "type", t, ln)) != -1) {
        if (t.size() != 2) throw LoadError("incorrect bead type definition");
        if (m.type_id.count(t[1])) throw LoadError("bead type redefinition");
        m.type_id[t[1]] = T++;
    }
    if (T <= 0) throw LoadError("no bead type definitions");
    m.s.types_num = T;
    const int np = (T * (T - 1)) / 2 + T;
    fn_pair blank{};
    blank.p1 = blank.p2 = -1;
    m.pairs.assign(np, blank);
    ln = 0;
    while ((ln = scan(L, "pair", t, ln)) != -1) {
        if (t.size() <= 4) throw LoadError("incomplete interaction type definition");
        if (!m.type_id.count(t[2]) || !m.type_id.count(t[3])) throw LoadError("pair type does not match any type");
        int a = m.type_id[t[2]], b = m.type_id[t[3]];
        if (a > b) std::swap(a, b);
        const int idx = T * a + b - (a * (a + 1)) / 2;
        fn_pair p{};
        if (t[1] == "lj") {
            if (t.size() == 7) {
                p.type = 1; p.p1 = a; p.p2 = b;
                p.c1 = stod_(t[4]); p.c2 = stod_(t[5]); p.c3 = stod_(t[6]);
                const double r2 = p.c3 * p.c3;
                double sigma2 = p.c1;
                const double eps = p.c2;
                sigma2 *= sigma2;
                const double fr2 = sigma2 / r2;
                const double fr6 = fr2 * fr2 * fr2;
                p.c4 = 48.0 * eps * fr6 * (fr6 - 0.5) / r2;
            }
        } else if (t[1] == "hs") {
            if (t.size() == 5) { p.type = 2; p.p1 = a; p.p2 = b; p.c1 = stod_(t[4]); }
        } else if (t[1] == "st") {
            if (t.size() == 8) {
                p.type = 3; p.p1 = a; p.p2 = b;
                p.c1 = stod_(t[5]);
                p.c2 = stod_(t[6]);
                p.c3 = stod_(t[7]) - p.c2;
                p.c4 = stod_(t[7]) * stod_(t[7]);
                p.c5 = stod_(t[4]);
                p.c6 = 2 * stod_(t[6]) - stod_(t[7]);
            }
        } else {
            throw LoadError("unknown interaction type");
        }
        m.pairs[idx] = p;
    }
    for (const fn_pair &p : m.pairs)
        if (p.p1 == -1 && p.p2 == -1) throw LoadError("pair list contains undefined interactions");
}

// SimulationManager.cpp:979-1091 LoadExportInfo.
void load_export(fn_model &m, const Lines &L)
{
    Tokens t;
    if (scan(L, "output_file", t, 0) <= 0) throw LoadError("output file name not specified");
    if (t.size() != 2) throw LoadError("output_file: incorrect number of command parameters");
    m.output_file = t[1];
    if (scan(L, "output_freq", t, 0) <= 0) throw LoadError("output_freq not specified");
    if (t.size() != 2) throw LoadError("output_freq: incorrect number of command parameters");
    m.dump = stoi_(t[1]);
    if (m.dump <= 0 || m.dump > m.run) throw LoadError("output_freq: incorrect value");
    std::vector<int> types;
    int ln = 0;
    while ((ln = scan(L, "output_type", t, ln)) != -1) {
        if (t.size() != 2) throw LoadError("output_type: incorrect number of command parameters");
        if (t[1] == "all") {
            types.clear();
            for (auto &kv : m.type_id) types.push_back(kv.second);
        } else {
            if (!m.type_id.count(t[1])) throw LoadError("output_type does not match any type");
            const int v = m.type_id[t[1]];
            if (std::find(types.begin(), types.end(), v) == types.end()) types.push_back(v);
        }
    }
    if (types.empty()) throw LoadError("list of exported bead types is empty");
    m.export_types = types;
    ln = 0;
    while ((ln = scan(L, "output_data", t, ln)) != -1) {
        if (t.size() != 2) throw LoadError("output_data: incorrect number of command parameters");
        if (t[1] == "position") m.export_position = true;
        else if (t[1] == "velocity") m.export_velocity = true;
        else if (t[1] == "force") m.export_force = true;
        else if (t[1] == "all") m.export_position = m.export_velocity = m.export_force = true;
        else throw LoadError("output_data: unknown data type");
    }
    if (!m.export_position && !m.export_velocity && !m.export_force) throw LoadError("export data type not specified");
}

void mass_coeffs(fn_bead_info &b, const fn_settings &s, double mass)
{
    b.mass = mass;
    if (b.mass > 0.0f) {
        b.mass_inv = 1.0f / b.mass;
        b.lambda_dt_mass_inv = s.lambda_dt * b.mass_inv;
        b.half_dt_mass_inv = s.half_dt * b.mass_inv;
        b.half_dt2_mass_inv = s.half_dt2 * b.mass_inv;
    } else {
        b.mass_inv = b.lambda_dt_mass_inv = b.half_dt_mass_inv = b.half_dt2_mass_inv = 0.0f;
    }
}

// SimulationManager.cpp:1372-1456 ReadBeadsData.
void read_beads(fn_model &m, const Lines &L)
{
    Tokens t;
    int ln = 0;
    while ((ln = scan(L, "BEAD", t, ln)) != -1) {
        if (t.size() != 18) throw LoadError("invalid Particle data definition at line " + std::to_string(ln));
        fn_bead b{};
        fn_bead_info q{};
        b.position.x = stod_(t[6]);
        b.position.y = stod_(t[7]);
        b.position.z = stod_(t[8]);
        const fn_vector3 &Lb = m.s.box_size;
        if (b.position.x > Lb.x || b.position.y > Lb.y || b.position.z > Lb.z || b.position.x < 0.0f ||
            b.position.y < 0.0f || b.position.z < 0.0f)
            throw LoadError("model geometry does not fit the simulation box size");
        b.velocity.x = stod_(t[9]);
        b.velocity.y = stod_(t[10]);
        b.velocity.z = stod_(t[11]);
        b.velocity_old.x = b.velocity.x;   // SimulationManager.cpp:1405-1407: all three take vx
        b.velocity_old.y = b.velocity.x;
        b.velocity_old.z = b.velocity.x;
        b.force.x = stod_(t[12]);
        b.force.y = stod_(t[13]);
        b.force.z = stod_(t[14]);
        b.force_old.x = b.force.x;
        b.force_old.y = b.force.y;
        b.force_old.z = b.force.z;
        b.image_position.x = stod_(t[15]);
        b.image_position.y = stod_(t[16]);
        b.image_position.z = stod_(t[17]);
        auto it = m.type_id.find(t[1]);
        if (it == m.type_id.end()) throw LoadError("bead type does not match any of defined types");
        q.type = it->second;
        q.tag = stoi_(t[2]);
        q.loc_id = stoi_(t[3]);
        q.radius = stod_(t[4]);
        q.radius_sq = q.radius * q.radius;
        q.skin_radius = q.radius + m.skin_radius_ratio;
        q.skin_radius_sq = q.skin_radius * q.skin_radius;
        mass_coeffs(q, m.s, stod_(t[5]));
        m.beads.push_back(b);
        m.info.push_back(q);
    }
    if (m.beads.empty()) throw LoadError("model data requires bead definition");
}

// SimulationManager.cpp:1501-1583 ReadSpringData.
void read_springs(fn_model &m, const Lines &L)
{
    Tokens t;
    int ln = 0;
    while ((ln = scan(L, "SPRING", t, ln)) != -1) {
        if (t.size() < 2) throw LoadError("invalid spring data definition");
        if (t[1] == "HR_FL" && t.size() != 9) throw LoadError("invalid spring data definition");
        if (t[1] == "HR" && t.size() < 7) throw LoadError("invalid spring data definition");
        fn_spring s{};
        if (t[1] == "HR") {
            s.type = 0;
        } else if (t[1] == "HR_FL") {
            s.type = 1;
            s.c1 = stoi_(t[7]);
            s.c2 = stoi_(t[8]);
        } else {
            throw LoadError("invalid spring data definition");
        }
        s.tag = stoi_(t[2]);
        s.p1 = stoi_(t[3]);
        s.p2 = stoi_(t[4]);
        s.rest_length = stod_(t[5]);
        if (s.rest_length < 0.0f) throw LoadError("wrong rest length value");
        s.stiffness = stod_(t[6]);
        if (s.stiffness < 0.0f) throw LoadError("wrong stiffness value");
        s.status = 1;
        m.springs.push_back(s);
    }
}

// SimulationManager.cpp:1591-1661 ReadAngleData.
void read_angles(fn_model &m, const Lines &L)
{
    Tokens t;
    int ln = 0;
    while ((ln = scan(L, "ANGLE", t, ln)) != -1) {
        if (t.size() != 10) throw LoadError("invalid angle data definition");
        fn_angle a{};
        if (t[1] == "HR") a.type = 0;
        else if (t[1] == "HR_FL") a.type = 1;
        else throw LoadError("invalid angle bond data definition");
        a.tag = stoi_(t[2]);
        a.b1 = stoi_(t[3]);
        a.b2 = stoi_(t[4]);
        a.b3 = stoi_(t[5]);
        a.s1 = stoi_(t[6]);
        a.s2 = stoi_(t[7]);
        a.angle = stof_(t[8]);
        a.c1 = stod_(t[9]);
        a.status = 1;
        m.angles.push_back(a);
    }
}

std::string resolve(const std::string &base, const std::string &p)
{
    if (p.empty() || p[0] == '/' || base.empty()) return p;
    return base + "/" + p;
}

// SimulationManager.cpp:1670-1823 ReadRestartFile.
void read_restart(fn_model &m, const Lines &L, const std::string &base)
{
    Tokens t;
    std::string name;
    int ln = 0;
    while ((ln = scan(L, "restart_file", t, ln)) != -1) {
        if (t.size() != 2) throw LoadError("restart_file: incorrect number of command parameters");
        name = t[1];
    }
    bool rp = false, rv = false, rf = false;
    ln = 0;
    while ((ln = scan(L, "restart_data", t, ln)) != -1) {
        if (t.size() != 2) throw LoadError("restart_data: incorrect number of command parameters");
        if (t[1] == "position") rp = true;
        else if (t[1] == "velocity") rv = true;
        else if (t[1] == "force") rf = true;
        else if (t[1] == "all") rp = rv = rf = true;
        else throw LoadError("restart_data: incorrect value");
    }
    if (name.empty()) return;
    Lines R;
    if (!read_lines(resolve(base, name), R)) throw LoadError("unable to open restart file " + name);
    if (R.empty()) throw LoadError("restart file is empty");
    const int n = (int)m.beads.size();
    int k = 0;
    for (const std::string &line : R) {
        t = tokens_of(line);
        if (t.size() != 13) continue;
        if (k >= n) throw LoadError("restart file has more beads than the model");
        auto it = m.type_id.find(t[0]);
        if (it == m.type_id.end()) throw LoadError("unable to recognise the bead type in the restart file");
        if (m.info[k].type != it->second) throw LoadError("restart bead type mismatch");
        fn_bead &b = m.beads[k];
        if (rp) {
            b.position.x = stod_(t[1]); b.position.y = stod_(t[2]); b.position.z = stod_(t[3]);
            b.image_position.x = stod_(t[4]); b.image_position.y = stod_(t[5]); b.image_position.z = stod_(t[6]);
        }
        if (rv) {
            b.velocity.x = stod_(t[7]); b.velocity.y = stod_(t[8]); b.velocity.z = stod_(t[9]);
            b.velocity_old.x = stod_(t[7]); b.velocity_old.y = stod_(t[8]); b.velocity_old.z = stod_(t[9]);
        }
        if (rf) {
            b.force.x = stod_(t[10]); b.force.y = stod_(t[11]); b.force.z = stod_(t[12]);
            b.force_old.x = b.force_old.y = b.force_old.z = 0.0;
        }
        ++k;
    }
    if (k != n) throw LoadError("incorrect number of beads in the restart file");
    m.restart = true;
}

// SimulationManager.cpp:189-415 LoadBeadModifications.
void bead_mods(fn_model &m, const Lines &L)
{
    Tokens t;
    int ln = 0;
    const int n = (int)m.beads.size();
    while ((ln = scan(L, "modif_bead", t, ln)) != -1) {
        if (t.size() <= 4) throw LoadError("modif_bead: incorrect number of arguments");
        int sel;   // 0 id, 1 tag, 2 type
        if (t[1] == "id") sel = 0;
        else if (t[1] == "tag") sel = 1;
        else if (t[1] == "type") sel = 2;
        else throw LoadError("modif_bead: unknown selection");
        int key;
        if (sel == 2) {
            auto it = m.type_id.find(t[2]);
            if (it == m.type_id.end()) throw LoadError("modif_bead: type does not match any type");
            key = it->second;
        } else {
            key = stoi_(t[2]);
        }
        auto match = [&](int i) {
            return (sel == 2 && m.info[i].type == key) || (sel == 1 && m.info[i].tag == key) || (sel == 0 && i == key);
        };
        if (t[3] == "mass") {
            if (t.size() != 5) throw LoadError("modif_bead mass: incorrect number of arguments");
            const double mass = stod_(t[4]);
            if (!(mass >= 0.0f)) throw LoadError("modif_bead mass: incorrect value");
            for (int i = 0; i < n; ++i)
                if (match(i)) mass_coeffs(m.info[i], m.s, mass);
        } else if (t[3] == "velocity") {
            if (t.size() != 7) throw LoadError("modif_bead velocity: incorrect number of arguments");
            fn_vector3 v{stod_(t[4]), stod_(t[5]), stod_(t[6]), 0.0};
            for (int i = 0; i < n; ++i)
                if (match(i)) { m.beads[i].velocity = v; m.beads[i].velocity_old = v; }
        } else if (t[3] == "skin_radius") {
            if (t.size() != 5) throw LoadError("modif_bead skin_radius: incorrect number of arguments");
            const double sk = stod_(t[4]);
            for (int i = 0; i < n; ++i)
                if (match(i)) { m.info[i].skin_radius = sk; m.info[i].skin_radius_sq = sk * sk; }
        } else {
            throw LoadError("modif_bead: unrecognized property");
        }
    }
}

// SimulationManager.cpp:423-636 LoadSpringModifications (which also requires
// the model to have springs: the spring buffer must exist, :429-432).
void spring_mods(fn_model &m, const Lines &L)
{
    if (m.springs.empty()) throw LoadError("spring buffer not initialized (the model has no springs)");
    Tokens t;
    int ln = 0;
    const int S = (int)m.springs.size();
    while ((ln = scan(L, "modif_spring", t, ln)) != -1) {
        if (t.size() <= 4) throw LoadError("modif_spring: incorrect number of arguments");
        int sel;
        if (t[1] == "id") sel = 0;
        else if (t[1] == "tag") sel = 1;
        else throw LoadError("modif_spring: unknown selection");
        const int key = stoi_(t[2]);
        double fn_spring::*field;
        if (t[3] == "c1") field = &fn_spring::c1;
        else if (t[3] == "c2") field = &fn_spring::c2;
        else if (t[3] == "rest_length") field = &fn_spring::rest_length;
        else if (t[3] == "stiffness") field = &fn_spring::stiffness;
        else if (t[3] == "damping") field = &fn_spring::damping;
        else throw LoadError("modif_spring: unrecognized property");
        if (t.size() != 5) {
            // the reference only logs this for stiffness and damping (SimulationManager.cpp:586,616)
            if (t[3] == "stiffness" || t[3] == "damping") continue;
            throw LoadError("modif_spring: incorrect number of arguments");
        }
        const double v = stod_(t[4]);
        if (!(v >= 0.0f)) throw LoadError("modif_spring: incorrect value");
        for (int s = 0; s < S; ++s)
            if ((sel == 1 && m.springs[s].tag == key) || (sel == 0 && s == key)) m.springs[s].*field = v;
    }
}

// SimulationManager::ReadInputFile (SimulationManager.cpp:1153-1171) and
// LoadModelData (SimulationManager.cpp:1315-1364).
void load_all_(fn_model &m, const std::string &path)
{
    Lines L;
    if (!read_lines(path, L)) throw LoadError("unable to open input file " + path);
    const size_t slash = path.find_last_of('/');
    const std::string base = slash == std::string::npos ? std::string() : path.substr(0, slash);
    load_global_force(m, L);
    load_solver(m, L);
    load_box(m, L);
    load_types_pairs(m, L);
    load_export(m, L);
    Tokens t;
    if (scan(L, "model_file", t, 0) <= 0) throw LoadError("model data file was not specified");
    if (t.size() != 2) throw LoadError("model_file: incorrect number of command parameters");
    Lines M;
    if (!read_lines(resolve(base, t[1]), M)) throw LoadError("unable to open model file " + t[1]);
    if (M.empty()) throw LoadError("model file is empty");
    read_beads(m, M);
    read_springs(m, M);
    read_angles(m, M);
    read_restart(m, L, base);
    bead_mods(m, L);
    spring_mods(m, L);
    m.output_file = resolve(base, m.output_file);
}

}  // namespace

void fnload::load_all(fn_model &m, const std::string &path) { load_all_(m, path); }
void fnload::set_error(const std::string &message) { g_model_error = message; }

extern "C" {

int fn_model_load(const char *input_path, fn_model **out)
{
    if (!input_path || !out) return FN_ERR_ARG;
    *out = nullptr;
    fn_model *m = new fn_model();
    try {
        load_all_(*m, input_path);
    } catch (const std::exception &e) {
        g_model_error = e.what();
        delete m;
        return FN_ERR_LOAD;
    }
    g_model_error.clear();
    *out = m;
    return FN_OK;
}

int fn_model_free(fn_model *m)
{
    delete m;
    return FN_OK;
}

const char *fn_model_error(void) { return g_model_error.c_str(); }

int fn_model_counts(const fn_model *m, int *nb, int *ns, int *na, int *np, int *nc)
{
    if (!m) return FN_ERR_ARG;
    if (nb) *nb = (int)m->beads.size();
    if (ns) *ns = (int)m->springs.size();
    if (na) *na = (int)m->angles.size();
    if (np) *np = (int)m->pairs.size();
    if (nc) *nc = (int)m->cellnh.size();
    return FN_OK;
}

int fn_model_params(const fn_model *m, int *run, int *output_freq, int *guf, int *min_steps, int *min_sub,
                    int *min_reset)
{
    if (!m) return FN_ERR_ARG;
    if (run) *run = m->run;
    if (output_freq) *output_freq = m->dump;
    if (guf) *guf = m->grid_update_freq;
    if (min_steps) *min_steps = m->min_steps;
    if (min_sub) *min_sub = m->min_sub;
    if (min_reset) *min_reset = m->min_reset ? 1 : 0;
    return FN_OK;
}

int fn_model_copy(const fn_model *m, int kind, void *dst)
{
    if (!m || !dst) return FN_ERR_ARG;
    switch (kind) {
    case 0: memcpy(dst, &m->s, sizeof(fn_settings)); break;
    case FN_BUF_BEAD: memcpy(dst, m->beads.data(), m->beads.size() * sizeof(fn_bead)); break;
    case FN_BUF_BEAD_INFO: memcpy(dst, m->info.data(), m->info.size() * sizeof(fn_bead_info)); break;
    case FN_BUF_SPRING: memcpy(dst, m->springs.data(), m->springs.size() * sizeof(fn_spring)); break;
    case FN_BUF_ANGLE: memcpy(dst, m->angles.data(), m->angles.size() * sizeof(fn_angle)); break;
    case FN_BUF_PAIR: memcpy(dst, m->pairs.data(), m->pairs.size() * sizeof(fn_pair)); break;
    case FN_BUF_GRID_CELL_NH: memcpy(dst, m->cellnh.data(), m->cellnh.size() * sizeof(fn_grid_cell_nh)); break;
    default: return FN_ERR_ARG;
    }
    return FN_OK;
}

}  // extern "C"
