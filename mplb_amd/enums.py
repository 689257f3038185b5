"""HiLeMMS enumerations, with the integer values the reference declares.

VertexType            Src/flowfield_host_device.h:21-37
VertexGeometryType    Src/flowfield_host_device.h:77-144
BoundarySurface       Src/boundary_host_device.h:7-29
BoundaryScheme        Src/boundary.h:71-84
VariableTypes         Src/model_host_device.h:6-19
CollisionType, BodyForceType, InitialType   Src/model.h:94-108
SchemeType            Src/scheme.h:63-68
The values travel unchanged through the C ABI (include/mplbhip.h).
"""
from enum import IntEnum


class VertexType(IntEnum):
    Fluid = 10
    Inlet = 11
    OutLet = 12
    MDPeriodic = 13
    Symmetry = 14
    FDPeriodic = 15
    Wall = 1000
    ImmersedSolid = -1
    ImmersedBoundary = -2
    VirtualBoundary = -3


class BoundarySurface(IntEnum):
    Left = 0
    Right = 1
    Top = 2
    Bottom = 3
    Front = 4
    Back = 5
    LeftBack = 50
    LeftFront = 40
    RightBack = 51
    RightFront = 41
    TopBack = 52
    TopFront = 42
    BottomBack = 53
    BottomFront = 43
    LeftTop = 20
    LeftBottom = 30
    RightTop = 21
    RightBottom = 31


class BoundaryScheme(IntEnum):
    KineticDiffuseWall = 11
    KineticSpelluarWall = 12
    ExtrapolPressure1ST = 16
    ExtrapolPressure2ND = 17
    MDPeriodic = 18
    FDPeriodic = 19
    BounceBack = 20
    FreeFlux = 21
    ZouHeVelocity = 22
    EQNNoSlip = 23
    EQMDiffuseRefl = 24
    # the reference spells the JSON string "EQMDiffuseREfl" (configuration.cpp:86)
    None_ = -1


class VariableTypes(IntEnum):
    Variable_Rho = 0
    Variable_U = 1
    Variable_V = 2
    Variable_W = 3
    Variable_T = 4
    Variable_Qx = 5
    Variable_Qy = 6
    Variable_Qz = 7
    Variable_U_Force = 8
    Variable_V_Force = 9
    Variable_W_Force = 10


class CollisionType(IntEnum):
    Collision_BGKIsothermal2nd = 0
    Collision_BGKIsothermal2nd_Swap = 10
    Collision_BGKThermal4th = 1
    Collision_BGKSWE4th = 2


class BodyForceType(IntEnum):
    BodyForce_None = 0
    BodyForce_1st = 1
    BodyForce_1st_Swap = 2
    BodyForce_None_Swap = 3


class InitialType(IntEnum):
    Initial_BGKFeq2nd = 1


class SchemeType(IntEnum):
    Scheme_E1st2nd = 1
    Scheme_I1st2nd = -1
    Scheme_StreamCollision = 10
    Scheme_StreamCollision_Swap = 11


# VertexGeometryType: faces, then edges/2D corners and 3D corners (inner)
VG_Fluid, VG_ImmersedSolid = 0, -1
VG_IP, VG_IM, VG_JP, VG_JM, VG_KP, VG_KM = 10, 11, 20, 21, 30, 31
