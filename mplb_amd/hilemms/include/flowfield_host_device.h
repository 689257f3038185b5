// flowfield_host_device.h -- node classification enums with the values of
// the reference (Src/flowfield_host_device.h:21-144).
#ifndef FLOWFIELD_HOST_DEVICE_H
#define FLOWFIELD_HOST_DEVICE_H
#ifndef OPS_FUN_PREFIX
#define OPS_FUN_PREFIX
#endif
enum class VertexType {
    Fluid = 10,
    Inlet = 11,
    OutLet = 12,
    MDPeriodic = 13,
    FDPeriodic = 15,
    Symmetry = 14,
    Wall = 1000,
    ImmersedSolid = -1,
    ImmersedBoundary = -2,
    VirtualBoundary = -3
};
enum VertexGeometryType {
    VG_Fluid = 0, VG_ImmersedSolid = -1,
    VG_IP = 10, VG_IM = 11, VG_JP = 20, VG_JM = 21, VG_KP = 30, VG_KM = 31,
    VG_IPKP_I = 10300, VG_IPKM_I = 10310, VG_IMKP_I = 11300, VG_IMKM_I = 11310,
    VG_JPKP_I = 20300, VG_JPKM_I = 20310, VG_JMKP_I = 21300, VG_JMKM_I = 21310,
    VG_IPJP_I = 10200, VG_IPJM_I = 10210, VG_IMJP_I = 11200, VG_IMJM_I = 11210,
    VG_IPJPKP_I = 1020300, VG_IPJPKM_I = 1020310, VG_IPJMKP_I = 1021300,
    VG_IPJMKM_I = 1021310, VG_IMJPKP_I = 1120300, VG_IMJPKM_I = 1120310,
    VG_IMJMKP_I = 1121300, VG_IMJMKM_I = 1121310,
    VG_IPKP_O = 10301, VG_IPKM_O = 10311, VG_IMKP_O = 11301, VG_IMKM_O = 11311,
    VG_JPKP_O = 20301, VG_JPKM_O = 20311, VG_JMKP_O = 21301, VG_JMKM_O = 21311,
    VG_IPJP_O = 10201, VG_IPJM_O = 10211, VG_IMJP_O = 11201, VG_IMJM_O = 11211,
    VG_IPJPKP_O = 1020301, VG_IPJPKM_O = 1020311, VG_IPJMKP_O = 1021301,
    VG_IPJMKM_O = 1021311, VG_IMJPKP_O = 1120301, VG_IMJPKM_O = 1120311,
    VG_IMJMKP_O = 1121301, VG_IMJMKM_O = 1121311
};
#ifdef OPS_3D
static inline OPS_FUN_PREFIX int SpaceDim() { return 3; }
#endif
#ifdef OPS_2D
static inline OPS_FUN_PREFIX int SpaceDim() { return 2; }
#endif
#endif
