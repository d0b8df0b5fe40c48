// Drives the engine through fibrenet::HipManager the way FibreNet's
// SimulationManager drives OpenclManager (buffers by id, arguments by index,
// kernels by name, a finish per step).  Built and run by tests/test_abi.py:
// without a GPU it must fail cleanly at Initialize; with one it runs a
// five-bead chain for ten steps of AdvancePosition and ResolveSprings and
// prints every bead's position, velocity and force, which the GPU test
// compares with the oracle's same ten steps.
#include <cstdio>
#include <string>
#include <vector>

#include "fibrenet_manager.hpp"

int main()
{
    fibrenet::HipManager m;
    std::string log;
    if (!m.Initialize([&](const std::string &s) { log += s + "\n"; }, FN_PREC_F64)) {
        std::printf("init failed\n");
        return 3;
    }
    fn_settings s{};
    s.types_num = 1; s.bc_type = 1; s.wall_contact_stiffness = 10.0; s.cutoff = 1.0; s.cutoff_sq = 1.0;
    s.box_unit_cell = 1.0; s.dt = 1e-3; s.half_dt = 5e-4; s.half_dt2 = 5e-7; s.lambda_dt = 5e-4;
    s.cells_x = s.cells_y = s.cells_z = 10; s.cells_xy = 100; s.half_box_sq = 25.0;
    s.grid_step = {1.0, 1.0, 1.0, 0.0}; s.box_size = {10.0, 10.0, 10.0, 0.0};
    std::vector<fn_bead> beads(5);
    std::vector<fn_bead_info> info(5);
    std::vector<fn_spring> springs(4);
    for (int i = 0; i < 5; ++i) {
        beads[i].position = beads[i].image_position = fn_vector3{2.0 + 1.1 * i, 5.0, 5.0, 0.0};
        info[i].tag = 0; info[i].loc_id = i; info[i].radius = 0.5; info[i].skin_radius = 1.0; info[i].mass = 1.0;
        info[i].mass_inv = 1.0; info[i].lambda_dt_mass_inv = 5e-4; info[i].half_dt_mass_inv = 5e-4;
        info[i].half_dt2_mass_inv = 5e-7;
    }
    for (int k = 0; k < 4; ++k) {
        springs[k].p1 = k; springs[k].p2 = k + 1; springs[k].status = 1; springs[k].rest_length = 1.0;
        springs[k].stiffness = 100.0;
    }
    std::vector<fn_pair> pairs(1);
    pairs[0].type = 2; pairs[0].c1 = 100.0;
    const int C = 1000;
    int head = m.CreateBuffer<int>(C, CL_MEM_READ_WRITE);
    int tail = m.CreateBuffer<fn_grid_cell>(C, CL_MEM_READ_WRITE);
    int pair = m.CreateBuffer<fn_pair>(pairs, CL_MEM_READ_ONLY);
    int bead = m.CreateBuffer<fn_bead>(beads, CL_MEM_READ_WRITE);
    int binf = m.CreateBuffer<fn_bead_info>(info, CL_MEM_READ_ONLY);
    int nhh = m.CreateBuffer<int>(5, CL_MEM_READ_WRITE);
    int nht = m.CreateBuffer<fn_particle_nh>(5, CL_MEM_READ_WRITE);
    int coef = m.CreateBuffer<fn_particle_nh>(5, CL_MEM_READ_WRITE);
    int spr = m.CreateBuffer<fn_spring>(springs, CL_MEM_READ_WRITE);
    if (beads.size() != 64 || springs.size() != 64) { std::printf("padding differs\n"); return 1; }
    std::vector<std::string> names = {"AdvancePosition", "ResetGrid", "ResetContactList", "FillGrid",
                                      "BuildContactList", "ResolveContacts", "AdvanceVelocity", "ResetBeadEnergy",
                                      "ResolveSprings", "ResolveAngles"};
    bool ok = m.RegisterKernel(names, "");
    int n = 5, sn = 4;
    ok &= m.SetKernelArg("AdvancePosition", 0, *m.GetBuffer(bead)) && m.SetKernelArg("AdvancePosition", 1, *m.GetBuffer(binf)) &&
          m.SetKernelArg("AdvancePosition", 2, s) && m.SetKernelArg("AdvancePosition", 3, n);
    ok &= m.SetKernelArg("ResolveSprings", 0, *m.GetBuffer(bead)) && m.SetKernelArg("ResolveSprings", 1, *m.GetBuffer(spr)) &&
          m.SetKernelArg("ResolveSprings", 2, sn);
    ok &= m.SetKernelArg("ResetGrid", 0, *m.GetBuffer(head)) && m.SetKernelArg("ResetGrid", 1, C);
    ok &= m.SetKernelArg("ResetContactList", 0, *m.GetBuffer(nhh)) && m.SetKernelArg("ResetContactList", 1, n);
    ok &= m.SetKernelArg("FillGrid", 0, *m.GetBuffer(bead)) && m.SetKernelArg("FillGrid", 1, *m.GetBuffer(head)) &&
          m.SetKernelArg("FillGrid", 2, *m.GetBuffer(tail)) && m.SetKernelArg("FillGrid", 3, s) && m.SetKernelArg("FillGrid", 4, n);
    int nhb = m.CreateBuffer<fn_grid_cell_nh>(C, CL_MEM_READ_ONLY);
    ok &= pair >= 0 && nht >= 0 && coef >= 0 && nhb >= 0;
    m.SetKernelGlobalWorkRange("AdvancePosition", m.GetBufferLength(bead));
    m.SetKernelGlobalWorkRange("ResolveSprings", m.GetBufferLength(spr));
    for (int step = 0; step < 10 && ok; ++step) {
        ok &= m.ExecuteKernel("AdvancePosition");
        ok &= m.ExecuteKernel("ResolveSprings");
        ok &= m.RunKernels();
    }
    std::vector<fn_bead> out(m.GetBufferLength(bead));
    ok &= m.ReadBuffer(bead, out.data());
    std::printf("%s x0=%.6f\n", ok ? "ok" : "failed", out[0].position.x);
    if (!ok) std::printf("%s", log.c_str());
    // every bead's position, velocity and force, for the comparison with the oracle
    for (int i = 0; i < 5 && ok; ++i) {
        const fn_bead &b = out[i];
        std::printf("bead %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", i, b.position.x, b.position.y,
                    b.position.z, b.velocity.x, b.velocity.y, b.velocity.z, b.force.x, b.force.y, b.force.z);
    }
    return ok ? 0 : 1;
}
