// configuration.cpp -- HiLeMMS JSON case files (keys and enum strings of
// the reference's Src/configuration.cpp:47-266).
#include "configuration.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

namespace {
Configuration config;
mplb_json::Value jsonConfig;

[[noreturn]] void bad(const std::string& what) {
    std::fprintf(stderr, "Error! %s\n", what.c_str());
    std::abort();
}

template <typename E>
E enum_of(const std::map<std::string, E>& table, const std::string& s, const char* kind) {
    auto it = table.find(s);
    if (it == table.end()) bad(std::string("unknown ") + kind + " \"" + s + "\"");
    return it->second;
}

const std::map<std::string, VariableTypes> kVar{
    {"Variable_Rho", Variable_Rho},     {"Variable_U", Variable_U},
    {"Variable_V", Variable_V},         {"Variable_W", Variable_W},
    {"Variable_T", Variable_T},         {"Variable_Qx", Variable_Qx},
    {"Variable_Qy", Variable_Qy},       {"Variable_Qz", Variable_Qz},
    {"Variable_U_Force", Variable_U_Force}, {"Variable_V_Force", Variable_V_Force},
    {"Variable_W_Force", Variable_W_Force}};
const std::map<std::string, VertexType> kVertex{
    {"Inlet", VertexType::Inlet},         {"OutLet", VertexType::OutLet},
    {"MDPeriodic", VertexType::MDPeriodic}, {"FDPeriodic", VertexType::FDPeriodic},
    {"Symmetry", VertexType::Symmetry},   {"Wall", VertexType::Wall},
    {"VirtualBoundary", VertexType::VirtualBoundary},
    {"ImmersedSolid", VertexType::ImmersedSolid},
    {"ImmersedBoundary", VertexType::ImmersedBoundary}};
const std::map<std::string, BoundaryScheme> kScheme{
    {"KineticDiffuseWall", BoundaryScheme::KineticDiffuseWall},
    {"ExtrapolPressure1ST", BoundaryScheme::ExtrapolPressure1ST},
    {"ExtrapolPressure2ND", BoundaryScheme::ExtrapolPressure2ND},
    {"MDPeriodic", BoundaryScheme::MDPeriodic},
    {"FDPeriodic", BoundaryScheme::FDPeriodic},
    {"BounceBack", BoundaryScheme::BounceBack},
    {"FreeFlux", BoundaryScheme::FreeFlux},
    {"ZouHeVelocity", BoundaryScheme::ZouHeVelocity},
    {"EQMDiffuseREfl", BoundaryScheme::EQMDiffuseRefl},
    {"None", BoundaryScheme::None}};
const std::map<std::string, CollisionType> kCollision{
    {"Collision_BGKIsothermal2nd", Collision_BGKIsothermal2nd},
    {"Collision_BGKIsothermal2nd_Swap", Collision_BGKIsothermal2nd_Swap},
    {"Collision_BGKThermal4th", Collision_BGKThermal4th},
    {"Collision_BGKSWE4th", Collision_BGKSWE4th}};
const std::map<std::string, BodyForceType> kForce{
    {"BodyForce_1st", BodyForce_1st}, {"BodyForce_None", BodyForce_None},
    {"BodyForce_1st_Swap", BodyForce_1st_Swap}, {"BodyForce_None_Swap", BodyForce_None_Swap}};
const std::map<std::string, BoundarySurface> kSurface{
    {"Left", BoundarySurface::Left},   {"Right", BoundarySurface::Right},
    {"Top", BoundarySurface::Top},     {"Bottom", BoundarySurface::Bottom},
#ifdef OPS_3D
    {"Front", BoundarySurface::Front}, {"Back", BoundarySurface::Back},
    {"LeftBack", BoundarySurface::LeftBack}, {"LeftFront", BoundarySurface::LeftFront},
    {"RightBack", BoundarySurface::RightBack}, {"RightFront", BoundarySurface::RightFront},
    {"TopBack", BoundarySurface::TopBack}, {"TopFront", BoundarySurface::TopFront},
    {"BottomBack", BoundarySurface::BottomBack},
    {"BottomFront", BoundarySurface::BottomFront},
#endif
    {"LeftTop", BoundarySurface::LeftTop}, {"LeftBottom", BoundarySurface::LeftBottom},
    {"RightTop", BoundarySurface::RightTop}, {"RightBottom", BoundarySurface::RightBottom}};
const std::map<std::string, InitialType> kInitial{{"Initial_BGKFeq2nd", Initial_BGKFeq2nd}};
const std::map<std::string, SchemeType> kSchemeType{
    {"Scheme_E1st2nd", Scheme_E1st2nd},
    {"Scheme_StreamCollision", Scheme_StreamCollision},
    {" Scheme_I1st2nd", Scheme_I1st2nd},
    {"Scheme_StreamCollision_Swap", Scheme_StreamCollision_Swap}};

const mplb_json::Value& need(const mplb_json::Value& obj, const std::string& key) {
    if (!obj.contains(key)) bad("Please supply " + key + " in the configuration!");
    if (obj[key].is_null()) bad("Please insert the " + key + " item into the configuration!");
    return obj[key];
}

void get(const mplb_json::Value& v, std::string& out) { out = v.str(); }
void get(const mplb_json::Value& v, int& out) { out = (int)v.number(); }
void get(const mplb_json::Value& v, SizeType& out) { out = (SizeType)v.number(); }
void get(const mplb_json::Value& v, double& out) { out = v.number(); }
void get(const mplb_json::Value& v, bool& out) { out = v.boolean(); }
void get(const mplb_json::Value& v, VariableTypes& out) { out = enum_of(kVar, v.str(), "variable"); }
void get(const mplb_json::Value& v, VertexType& out) { out = enum_of(kVertex, v.str(), "vertex type"); }
void get(const mplb_json::Value& v, BoundaryScheme& out) { out = enum_of(kScheme, v.str(), "boundary scheme"); }
void get(const mplb_json::Value& v, CollisionType& out) { out = enum_of(kCollision, v.str(), "collision"); }
void get(const mplb_json::Value& v, BodyForceType& out) { out = enum_of(kForce, v.str(), "body force"); }
void get(const mplb_json::Value& v, BoundarySurface& out) { out = enum_of(kSurface, v.str(), "surface"); }
void get(const mplb_json::Value& v, InitialType& out) { out = enum_of(kInitial, v.str(), "initial type"); }
void get(const mplb_json::Value& v, SchemeType& out) { out = enum_of(kSchemeType, v.str(), "scheme"); }
template <typename T>
void get(const mplb_json::Value& v, std::vector<T>& out) {
    out.clear();
    for (const auto& e : v.array()) {
        T x;
        get(e, x);
        out.push_back(x);
    }
}
template <typename T>
void Query(T& value, const std::string& key) { get(need(jsonConfig, key), value); }
template <typename T>
void Check(T& value, const std::string& key) {
    if (!jsonConfig.contains(key) || jsonConfig[key].is_null()) {
        std::printf("Warning! %s is not defined in the configuration!\n", key.c_str());
        return;
    }
    get(jsonConfig[key], value);
}
template <typename T>
void Query(T& value, const std::string& k0, const std::string& k1) {
    get(need(need(jsonConfig, k0), k1), value);
}

int boundary_condition_count() {
    int n = 0;
    while (jsonConfig.contains("BoundaryCondition" + std::to_string(n)) &&
           !jsonConfig["BoundaryCondition" + std::to_string(n)].is_null())
        n++;
    return n;
}

// ParseJson, configuration.cpp:152-223
void ParseJson() {
    Query(config.caseName, "CaseName");
    Query(config.spaceDim, "SpaceDim");
    Query(config.compoNames, "CompoNames");
    Query(config.lattNames, "LatticeName");
    Query(config.compoIds, "CompoIds");
    Query(config.tauRef, "TauRef");
    Query(config.macroVarNames, "MacroVarNames");
    Query(config.macroVarIds, "MacroVarIds");
    Query(config.macroCompoIds, "MacroCompoIds");
    Query(config.macroVarTypes, "MacroVarTypes");
    Query(config.CollisionTypes, "CollisionType");
    Query(config.CollisionCompoIds, "CollisionCompoIds");
    Query(config.initialTypes, "InitialType");
    Query(config.initialConditionCompoId, "InitialCompoIds");
    Query(config.bodyForceCompoIds, "BodyForceCompoId");
    Query(config.bodyForceTypes, "BodyForceType");
    Query(config.schemeType, "SchemeType");
    Query(config.blockIds, "BlockIds");
    Query(config.blockNames, "BlockNames");
    Query(config.blockSize, "BlockSize");
    Check(config.fromBlockIds, "FromBlockIds");
    Check(config.toBlockIds, "ToBlockIds");
    Check(config.fromBoundarySurface, "FromBoundarySurface");
    Check(config.toBoundarySurface, "ToBoundarySurface");
    Check(config.blockConnectionType, "BlockConnectionType");
    for (const auto id : config.blockIds) {
        std::vector<Real> pos;
        Query(pos, "StartPos", std::to_string(id));
        config.startPos.emplace(id, pos);
    }
    Query(config.checkPeriod, "CheckPeriod");
    Query(config.meshSize, "MeshSize");
    const int nbc = boundary_condition_count();
    config.blockBoundaryConfig.resize(nbc);
    for (int i = 0; i < nbc; i++) {
        const std::string k{"BoundaryCondition" + std::to_string(i)};
        BlockBoundary& b = config.blockBoundaryConfig[i];
        Query(b.blockIndex, k, "BlockIndex");
        Query(b.componentID, k, "ComponentId");
        Query(b.boundarySurface, k, "BoundarySurface");
        Query(b.boundaryScheme, k, "BoundaryScheme");
        Query(b.givenVars, k, "GivenVars");
        Query(b.boundaryType, k, "BoundaryType");
        Query(b.macroVarTypesatBoundary, k, "MacroVarTypesatBoundary");
    }
    Query(config.currentTimeStep, "CurrentTimeStep");
    Query(config.transient, "Transient");
    if (config.transient) Query(config.timeStepsToRun, "TimeStepsToRun");
    else Query(config.convergenceCriteria, "ConvergenceCriteria");
}
}  // namespace

const Configuration& Config() { return config; }
const mplb_json::Value& JsonConfig() { return jsonConfig; }

void ReadConfiguration(std::string& configFileName) {
    std::ifstream in(configFileName);
    if (!in.is_open()) bad("Cannot open the configuration file " + configFileName);
    std::stringstream ss;
    ss << in.rdbuf();
    try {
        jsonConfig = mplb_json::Value::parse(ss.str());
    } catch (const mplb_json::ParseError& e) {
        bad(e.what());
    }
    config = Configuration();
    ParseJson();
}

void GetConfigFileFromCmd(bool& findConfig, std::string& fileName, const int argc,
                          const char** argv) {
    findConfig = false;
    for (int i = 1; i < argc; i++) {
        const std::string arg{argv[i]};
        if (arg.find("Config=") == 0) {
            findConfig = true;
            fileName = arg.substr(7);
            break;
        }
    }
}
