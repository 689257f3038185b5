// configuration.h -- JSON case configuration (reference
// Src/configuration.h:40-191).  The JSON reader is this library's own
// (json_value.h); enum strings follow the reference, including
// "EQMDiffuseREfl" and " Scheme_I1st2nd" (configuration.cpp:86, :135).
#ifndef CONFIGURATION_H
#define CONFIGURATION_H
#include <map>
#include <string>
#include <vector>

#include "boundary.h"
#include "flowfield_host_device.h"
#include "json_value.h"
#include "model.h"
#include "model_host_device.h"
#include "scheme.h"
#include "type.h"

struct Configuration {
    std::string caseName;
    SizeType spaceDim{3};
    std::vector<std::string> compoNames;
    std::vector<int> compoIds;
    std::vector<std::string> lattNames;
    std::vector<VariableTypes> macroVarTypes;
    std::vector<std::string> macroVarNames;
    std::vector<int> macroVarIds;
    std::vector<int> macroCompoIds;
    std::vector<CollisionType> CollisionTypes;
    std::vector<int> CollisionCompoIds;
    std::vector<BodyForceType> bodyForceTypes;
    std::vector<SizeType> bodyForceCompoIds;
    std::vector<InitialType> initialTypes;
    std::vector<int> initialConditionCompoId;
    SchemeType schemeType{Scheme_StreamCollision};
    std::vector<std::string> blockNames;
    std::vector<int> blockIds;
    std::vector<int> blockSize;
    std::vector<int> fromBlockIds;
    std::vector<int> toBlockIds;
    std::vector<BoundarySurface> fromBoundarySurface;
    std::vector<BoundarySurface> toBoundarySurface;
    std::vector<VertexType> blockConnectionType;
    std::map<int, std::vector<Real>> startPos;
    Real meshSize{0};
    std::vector<Real> tauRef;
    bool transient{true};
    Real convergenceCriteria{-1};
    SizeType timeStepsToRun{0};
    SizeType currentTimeStep{0};
    SizeType checkPeriod{1000};
    std::vector<BlockBoundary> blockBoundaryConfig;
};

void ReadConfiguration(std::string& configFileName);
const Configuration& Config();
const mplb_json::Value& JsonConfig();
void GetConfigFileFromCmd(bool& findConfig, std::string& fileName, const int argc,
                          const char** argv);
#endif
