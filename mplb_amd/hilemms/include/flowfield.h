// flowfield.h -- global fields, blocks, case definition, residual and the
// app hooks (reference Src/flowfield.h:44-136).
#ifndef FLOWFIELD_H
#define FLOWFIELD_H
#include <map>
#include <string>

#include "block.h"
#include "field.h"
#include "type.h"

const BlockGroup& g_Block();
RealField& g_f();
RealField& g_fStage();
RealFieldGroup& g_MacroVars();
RealFieldGroup& g_MacroVarsCopy();
RealFieldGroup& g_MacroBodyforce();
RealField& g_CoordinateXYZ();
IntFieldGroup& g_NodeType();
IntField& g_GeometryProperty();
Real TimeStep();
const Real* pTimeStep();
const std::string& CaseName();
Real TotalMeshSize();
void SetTimeStep(Real dt);
std::map<int, Real>& g_ResidualError();

void WriteFlowfieldToHdf5(const SizeType timeStep);
void WriteDistributionsToHdf5(const SizeType timeStep);
void WriteNodePropertyToHdf5(const SizeType timeStep);
void Partition();
void PrepareFlowField();
void DefineCase(const std::string& caseName, const int spaceDim,
                const bool transient = false);
Real GetMaximumResidual(const SizeType checkPeriod);
void DefineBlocks(const std::vector<int>& blockIds,
                  const std::vector<std::string>& blockNames,
                  const std::vector<int>& blockSizes, const Real meshSize,
                  const std::map<int, std::vector<Real>>& startPos);
bool IsTransient();
void CalcResidualError();
void DispResidualError(const int iter, const SizeType checkPeriod);
void CopyCurrentMacroVar();
// app-defined hooks (flowfield.h:114-115 of the reference)
void UpdateMacroscopicBodyForce(const Real time);
void SetInitialMacrosVars();
void DefineBlockConnection(const std::vector<int>& fromBlock,
                           const std::vector<BoundarySurface>& fromSurface,
                           const std::vector<int>& toBlock,
                           const std::vector<BoundarySurface>& toSurface,
                           const std::vector<VertexType>& connectionType);
void TransferHalos();
#endif
