// case_driver.cpp -- runs any HiLeMMS JSON case file (Config=<file>) through
// the same call sequence as the reference Apps' simulate(config), on one
// GPU.  Build with DIM=2 or DIM=3.  Initial state: rest, density 1.
//   case_driver Config=<file.json> [Dump=<file>] [Force=<g0>] [ForceSchedule=<g0>]
//               [Temperature=<T0>]
// Force gives the body-force hook a constant acceleration g0 along x;
// ForceSchedule a time-dependent one, g0 (1 + floor(k / 7) mod 3) at step k
// (time k dt); either is written into g_MacroBodyforce as an App's hook
// would.  Temperature gives a Variable_T macro, when the case defines one,
// the uniform value T0 (default 1).
#include <cmath>
#include <iostream>
#include <string>

#include "app_common.h"
#include "mplb.h"
#include "ops_seq_v2.h"

#ifdef OPS_3D
void KerRestInitial(ACC<Real>& rho, ACC<Real>& u, ACC<Real>& v, ACC<Real>& w) {
    rho(0, 0, 0) = 1;
    u(0, 0, 0) = 0;
    v(0, 0, 0) = 0;
    w(0, 0, 0) = 0;
}
#else
void KerRestInitial(ACC<Real>& rho, ACC<Real>& u, ACC<Real>& v) {
    rho(0, 0) = 1;
    u(0, 0) = 0;
    v(0, 0) = 0;
}
#endif

static double g_temperature = 1.0;

#ifdef OPS_3D
void KerSetTemperature(ACC<Real>& T, const Real* val) { T(0, 0, 0) = val[0]; }
#else
void KerSetTemperature(ACC<Real>& T, const Real* val) { T(0, 0) = val[0]; }
#endif

void SetInitialMacrosVars() {
    for (auto idBlock : g_Block()) {
        Block& block{idBlock.second};
        std::vector<int> rng(block.WholeRange().begin(), block.WholeRange().end());
        for (auto& idCompo : g_Components()) {
            const Component& c{idCompo.second};
            // Variable_T (read by Collision_BGKThermal4th): Temperature=<T0>
            if (c.macroVars.count(Variable_T))
                ops_par_loop(KerSetTemperature, "KerSetTemperature", block.Get(), SpaceDim(),
                             rng.data(),
                             ops_arg_dat(g_MacroVars().at(c.macroVars.at(Variable_T).id)
                                             .at(block.ID()),
                                         1, LOCALSTENCIL, "double", OPS_RW),
                             ops_arg_gbl(&g_temperature, 1, "double", OPS_READ));
            const int rhoId{c.macroVars.at(Variable_Rho).id};
            ops_par_loop(KerRestInitial, "KerRestInitial", block.Get(), SpaceDim(), rng.data(),
                         ops_arg_dat(g_MacroVars().at(rhoId).at(block.ID()), 1, LOCALSTENCIL,
                                     "double", OPS_RW),
                         ops_arg_dat(g_MacroVars().at(c.uId).at(block.ID()), 1, LOCALSTENCIL,
                                     "double", OPS_RW),
                         ops_arg_dat(g_MacroVars().at(c.vId).at(block.ID()), 1, LOCALSTENCIL,
                                     "double", OPS_RW)
#ifdef OPS_3D
                             ,
                         ops_arg_dat(g_MacroVars().at(c.wId).at(block.ID()), 1, LOCALSTENCIL,
                                     "double", OPS_RW)
#endif
            );
        }
    }
}
static double g_force_amp = 0.0;
static bool g_force_schedule = false;

#ifdef OPS_3D
void KerSetAccel(ACC<Real>& g, const Real* val) {
    g(0, 0, 0, 0) = val[0];
    g(1, 0, 0, 0) = val[1];
    g(2, 0, 0, 0) = val[2];
}
#else
void KerSetAccel(ACC<Real>& g, const Real* val) {
    g(0, 0, 0) = val[0];
    g(1, 0, 0) = val[1];
}
#endif

void UpdateMacroscopicBodyForce(const Real time) {
    if (g_force_amp == 0.0) return;
    const long long k = std::llround(time / TimeStep());
    const Real val[3] = {g_force_schedule ? g_force_amp * (1 + (k / 7) % 3) : g_force_amp, 0.0,
                         0.0};
    for (auto idBlock : g_Block()) {
        Block& block{idBlock.second};
        std::vector<int> rng(block.WholeRange().begin(), block.WholeRange().end());
        for (auto& idForce : g_MacroBodyforce())
            ops_par_loop(KerSetAccel, "KerSetAccel", block.Get(), SpaceDim(), rng.data(),
                         ops_arg_dat(idForce.second.at(block.ID()), SpaceDim(), LOCALSTENCIL,
                                     "double", OPS_RW),
                         ops_arg_gbl(val, 3, "double", OPS_READ));
    }
}

static void simulate(const Configuration& config) {
    DefineCase(config.caseName, config.spaceDim, config.transient);
    DefineBlocks(config.blockIds, config.blockNames, config.blockSize, config.meshSize,
                 config.startPos);
    if (!config.fromBlockIds.empty())
        DefineBlockConnection(config.fromBlockIds, config.fromBoundarySurface,
                              config.toBlockIds, config.toBoundarySurface,
                              config.blockConnectionType);
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
    if (config.currentTimeStep == 0) {
        SetInitialMacrosVars();
#ifdef OPS_3D
        PreDefinedInitialCondition3D();
#else
        PreDefinedInitialCondition();
#endif
    }
    SetTimeStep(config.meshSize / SoundSpeed());
    // the cycle an App passes for its scheme (Apps/3DCavity-Swap uses
    // SwapStreamCollision, lbm3d_cavity_swap.cpp:101-107)
    void (*cycle)(Real) = config.schemeType == Scheme_StreamCollision_Swap
                              ? &SwapStreamCollision
                              : &StreamCollision;
    if (config.transient)
        Iterate(cycle, config.timeStepsToRun, config.checkPeriod, config.currentTimeStep);
    else
        Iterate(cycle, config.convergenceCriteria, config.checkPeriod, config.currentTimeStep);
}

int main(int argc, const char** argv) {
    ops_init(argc, argv, 1);
    const auto o = AppOptions(argc, argv);
    bool found = false;
    std::string cfg;
    GetConfigFileFromCmd(found, cfg, argc, argv);
    if (!found) {
        ops_printf("usage: case_driver Config=<case.json> [Dump=<file>]\n");
        return 2;
    }
    double c0, e0, c1, e1;
    ops_timers(&c0, &e0);
    ReadConfiguration(cfg);
    g_force_amp = OptD(o, "Force", 0.0);
    g_temperature = OptD(o, "Temperature", 1.0);
    if (o.count("ForceSchedule")) {
        g_force_amp = OptD(o, "ForceSchedule", 0.0);
        g_force_schedule = true;
    }
    if (o.count("DryRun")) {
        // parsed case, no device: blocks, boundaries, steps
        const Configuration& c = Config();
        std::printf("case=%s dim=%zu blocks=%zu bcs=%zu links=%zu tau=%.17g mesh=%.17g "
                    "check=%zu steps=%zu transient=%d\n",
                    c.caseName.c_str(), c.spaceDim, c.blockIds.size(),
                    c.blockBoundaryConfig.size(), c.fromBlockIds.size(), c.tauRef.at(0),
                    c.meshSize, c.checkPeriod, c.timeStepsToRun, (int)c.transient);
        for (const auto& bc : c.blockBoundaryConfig)
            std::printf("bc block=%d surface=%d scheme=%d type=%d nvars=%zu\n", bc.blockIndex,
                        (int)bc.boundarySurface, (int)bc.boundaryScheme, (int)bc.boundaryType,
                        bc.givenVars.size());
        return 0;
    }
    simulate(Config());
    ops_timers(&c1, &e1);
    ops_printf("\nTotal Wall time %lf\n", e1 - e0);
    if (o.count("Dump")) DumpMacroVars(o.at("Dump"));
    ops_exit();
    return 0;
}
