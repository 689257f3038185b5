// cavity3d.cpp -- 3D lid-driven cavity through the HiLeMMS case API.
//   cavity3d [N=33] [Re=400] [Lid=0.05] [Steps=1000] [Check=1000]
//            [Noise=0|1] [Criteria=<residual>] [Dump=<file>] [Config=<json>]
// Lattice units: meshSize = 1/(N-1), dt = meshSize/CS, tau_L = 0.5 + 3 Lid N/Re,
// TauRef = (tau_L - 1/2) dt, lid code velocity CS*Lid.  Walls are
// EQMDiffuseRefl, defined Left, Right, Top, Bottom, Front, Back.
#include <cmath>
#include <iostream>
#include <string>

#include "app_common.h"
#include "mplb.h"
#include "ops_seq_v2.h"
#include "cavity3d_init.inc"

void SetInitialMacrosVars() {
    for (auto idBlock : g_Block()) {
        Block& block{idBlock.second};
        std::vector<int> rng(block.WholeRange().begin(), block.WholeRange().end());
        g_nx = block.Size()[0];
        g_ny = block.Size()[1];
        for (auto& idCompo : g_Components()) {
            const Component& c{idCompo.second};
            const int rhoId{c.macroVars.at(Variable_Rho).id};
            ops_par_loop(KerCavityInitial3D, "KerCavityInitial3D", block.Get(), SpaceDim(),
                         rng.data(),
                         ops_arg_dat(g_MacroVars().at(rhoId).at(block.ID()), 1,
                                     LOCALSTENCIL, "double", OPS_RW),
                         ops_arg_dat(g_MacroVars().at(c.uId).at(block.ID()), 1,
                                     LOCALSTENCIL, "double", OPS_RW),
                         ops_arg_dat(g_MacroVars().at(c.vId).at(block.ID()), 1,
                                     LOCALSTENCIL, "double", OPS_RW),
                         ops_arg_dat(g_MacroVars().at(c.wId).at(block.ID()), 1,
                                     LOCALSTENCIL, "double", OPS_RW),
                         ops_arg_idx());
        }
    }
}
void UpdateMacroscopicBodyForce(const Real time) {}

static void simulate(const std::map<std::string, std::string>& o) {
    const int n = (int)OptD(o, "N", 33);
    const double re = OptD(o, "Re", 400.0), lid = OptD(o, "Lid", 0.05);
    g_noiseAmp = OptD(o, "Noise", 0) != 0 ? 1e-4 : 0.0;
    DefineCase("3D_cavity", 3);
    const Real mesh = 1.0 / (n - 1);
    DefineBlocks({0}, {"Cavity"}, {n, n, n}, mesh, {{0, {0.0, 0.0, 0.0}}});
    const Real dt = mesh / std::sqrt(3.0);
    const Real tauL = 0.5 + 3.0 * lid * n / re;
    DefineComponents({"Fluid"}, {0}, {"d3q19"}, {(tauL - 0.5) * dt});
    DefineMacroVars({Variable_Rho, Variable_U, Variable_V, Variable_W},
                    {"rho", "u", "v", "w"}, {0, 1, 2, 3}, {0, 0, 0, 0});
    DefineCollision({Collision_BGKIsothermal2nd}, {0});
    DefineBodyForce({BodyForce_None}, {0});
    DefineScheme(Scheme_StreamCollision);
    const std::vector<VariableTypes> uvw{Variable_U, Variable_V, Variable_W};
    const std::vector<Real> still{0, 0, 0}, moving{std::sqrt(3.0) * lid, 0, 0};
    for (auto s : {BoundarySurface::Left, BoundarySurface::Right, BoundarySurface::Top,
                   BoundarySurface::Bottom, BoundarySurface::Front, BoundarySurface::Back})
        DefineBlockBoundary(0, 0, s, BoundaryScheme::EQMDiffuseRefl, uvw,
                            s == BoundarySurface::Top ? moving : still);
    DefineInitialCondition({Initial_BGKFeq2nd}, {0});
    Partition();
    SetInitialMacrosVars();
    PreDefinedInitialCondition3D();
    SetTimeStep(dt);
    const SizeType check = (SizeType)OptD(o, "Check", 1000);
    if (o.count("Criteria"))
        Iterate(StreamCollision, (Real)OptD(o, "Criteria", 1e-8), check);
    else
        Iterate(StreamCollision, (SizeType)OptD(o, "Steps", 1000), check);
}

static void simulate(const Configuration& config) {
    DefineCase(config.caseName, config.spaceDim, config.transient);
    DefineBlocks(config.blockIds, config.blockNames, config.blockSize, config.meshSize,
                 config.startPos);
    DefineComponents(config.compoNames, config.compoIds, config.lattNames, config.tauRef,
                     config.currentTimeStep);
    DefineMacroVars(config.macroVarTypes, config.macroVarNames, config.macroVarIds,
                    config.macroCompoIds, config.currentTimeStep);
    DefineCollision(config.CollisionTypes, config.CollisionCompoIds);
    DefineBodyForce(config.bodyForceTypes, config.bodyForceCompoIds);
    DefineScheme(config.schemeType);
    DefineInitialCondition(config.initialTypes, config.initialConditionCompoId);
    for (auto& bc : config.blockBoundaryConfig)
        DefineBlockBoundary(bc.blockIndex, bc.componentID, bc.boundarySurface,
                            bc.boundaryScheme, bc.macroVarTypesatBoundary, bc.givenVars,
                            bc.boundaryType);
    Partition();
    SetInitialMacrosVars();
    PreDefinedInitialCondition3D();
    SetTimeStep(config.meshSize / SoundSpeed());
    if (config.transient)
        Iterate(config.timeStepsToRun, config.checkPeriod, config.currentTimeStep);
    else
        Iterate(config.convergenceCriteria, config.checkPeriod, config.currentTimeStep);
}

int main(int argc, const char** argv) {
    ops_init(argc, argv, 1);
    const auto o = AppOptions(argc, argv);
    bool found = false;
    std::string cfg;
    GetConfigFileFromCmd(found, cfg, argc, argv);
    double c0, e0, c1, e1;
    ops_timers(&c0, &e0);
    if (found) {
        ReadConfiguration(cfg);
        simulate(Config());
    } else {
        simulate(o);
    }
    ops_timers(&c1, &e1);
    ops_printf("\nTotal Wall time %lf\n", e1 - e0);
    if (o.count("Dump")) DumpMacroVars(o.at("Dump"));
    ops_exit();
    return 0;
}
