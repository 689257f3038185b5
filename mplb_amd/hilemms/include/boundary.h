// boundary.h -- block boundary registry (reference Src/boundary.h:40-126).
// The boundary kernels themselves run inside libmplbhip's fused step.
#ifndef BOUNDARY_H
#define BOUNDARY_H
#include <vector>

#include "block.h"
#include "boundary_host_device.h"
#include "flowfield_host_device.h"
#include "model.h"
#include "model_host_device.h"

const std::vector<BoundarySurface> AllBoundarySurface{
    BoundarySurface::Left,      BoundarySurface::Right,    BoundarySurface::Top,
    BoundarySurface::Bottom,
#ifdef OPS_3D
    BoundarySurface::Front,     BoundarySurface::Back,     BoundarySurface::LeftBack,
    BoundarySurface::LeftFront, BoundarySurface::RightBack, BoundarySurface::RightFront,
    BoundarySurface::TopBack,   BoundarySurface::TopFront, BoundarySurface::BottomBack,
    BoundarySurface::BottomFront,
#endif
    BoundarySurface::LeftTop,   BoundarySurface::LeftBottom, BoundarySurface::RightTop,
    BoundarySurface::RightBottom};

enum class BoundaryScheme {
    KineticDiffuseWall = 11,
    KineticSpelluarWall = 12,
    ExtrapolPressure1ST = 16,
    ExtrapolPressure2ND = 17,
    MDPeriodic = 18,
    FDPeriodic = 19,
    BounceBack = 20,
    FreeFlux = 21,
    ZouHeVelocity = 22,
    EQNNoSlip = 23,
    EQMDiffuseRefl = 24,
    None = -1
};

struct BlockBoundary {
    int blockIndex;
    int componentID;
    std::vector<Real> givenVars;
    BoundarySurface boundarySurface;
    BoundaryScheme boundaryScheme;
    std::vector<VariableTypes> macroVarTypesatBoundary;
    VertexType boundaryType;
};

int BoundaryHaloNum();
void SetBoundaryHaloNum(const int boundaryhaloNum);
void DefineBlockBoundary(int blockIndex, int componentID, BoundarySurface boundarySurface,
                         BoundaryScheme boundaryScheme,
                         const std::vector<VariableTypes>& macroVarTypes,
                         const std::vector<Real>& macroVarValues,
                         const VertexType boundaryType = VertexType::Wall);
void DefineBlockBoundary(int blockIndex, int componentID, BoundarySurface boundarySurface,
                         const VertexType boundaryType = VertexType::VirtualBoundary);
const std::vector<BlockBoundary>& BlockBoundaries();
#ifdef OPS_3D
void ImplementBoundary3D();
#endif
#ifdef OPS_2D
void ImplementBoundary();
#endif
#endif
