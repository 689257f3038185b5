// model.h -- components, macroscopic variables, collision / body force /
// initial-condition kinds (reference Src/model.h:48-177).  The lattice
// tables are exported for Apps that read them on the host.
#ifndef MODEL_H
#define MODEL_H
#include <list>
#include <map>
#include <string>
#include <vector>

#include "model_host_device.h"
#include "type.h"

extern int NUMXI;
extern int LATTDIM;
extern Real CS;
extern Real* XI;
extern Real XIMAXVALUE;
extern Real* WEIGHTS;
extern int NUMCOMPONENTS;
extern int* OPP;

enum CollisionType {
    Collision_BGKIsothermal2nd = 0,
    Collision_BGKIsothermal2nd_Swap = 10,
    Collision_BGKThermal4th = 1,
    Collision_BGKSWE4th = 2,
};
enum BodyForceType {
    BodyForce_1st = 1,
    BodyForce_None = 0,
    BodyForce_1st_Swap = 2,
    BodyForce_None_Swap = 3,
};
enum InitialType { Initial_BGKFeq2nd = 1 };

struct MacroVariable {
    std::string name;
    int id{0};
    VariableTypes type;
};
struct Component {
    std::string name;
    int id{0};
    std::string latticeName;
    CollisionType collisionType{Collision_BGKIsothermal2nd};
    BodyForceType bodyForceType{BodyForce_None};
    InitialType initialType{Initial_BGKFeq2nd};
    std::map<VariableTypes, MacroVariable> macroVars;
    int index[2]{0, 0};
    int uId{0}, vId{0};
#ifdef OPS_3D
    int wId{0};
#endif
    Real tauRef{0};
};

const std::map<int, Component>& g_Components();
inline int ComponentNum() { return NUMCOMPONENTS; }
inline int SizeF() { return NUMXI; }
inline Real SoundSpeed() { return CS; }
inline Real MaximumSpeed() { return XIMAXVALUE; }
void DestroyModel();

void DefineComponents(const std::vector<std::string>& compoNames,
                      const std::vector<int>& compoId,
                      const std::vector<std::string>& lattNames,
                      const std::vector<Real> tauRef, const SizeType timeStep = 0);
void DefineMacroVars(std::vector<VariableTypes> types, std::vector<std::string> names,
                     std::vector<int> varId, std::vector<int> compoId,
                     const SizeType timeStep = 0);
void DefineCollision(std::vector<CollisionType> types, std::vector<int> compoId);
void DefineBodyForce(std::vector<BodyForceType> types, std::vector<SizeType> compoId);
void DefineInitialCondition(std::vector<InitialType> types, std::vector<int> compoId);
#ifdef OPS_3D
void UpdateMacroVars3D();
void PreDefinedBodyForce3D();
void PreDefinedInitialCondition3D();
void PreDefinedCollision3D();
#endif
#ifdef OPS_2D
void UpdateMacroVars();
void PreDefinedBodyForce();
void PreDefinedInitialCondition();
void PreDefinedCollision();
#endif
#endif
