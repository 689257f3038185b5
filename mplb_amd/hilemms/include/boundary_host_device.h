// boundary_host_device.h -- boundary surfaces and velocity classes with
// the values of the reference (Src/boundary_host_device.h:7-39).  The
// classification itself runs inside libmplbhip (region programs).
#ifndef BOUNDARY_HOST_DEVICE_H
#define BOUNDARY_HOST_DEVICE_H
enum class BoundarySurface {
    Left = 0,
    Right = 1,
    Top = 2,
    Bottom = 3,
#ifdef OPS_3D
    Front = 4,
    Back = 5,
    LeftBack = 50,
    LeftFront = 40,
    RightBack = 51,
    RightFront = 41,
    TopBack = 52,
    TopFront = 42,
    BottomBack = 53,
    BottomFront = 43,
#endif
    LeftTop = 20,
    LeftBottom = 30,
    RightTop = 21,
    RightBottom = 31
};
enum BndryDvType {
    BndryDv_Incoming = 1,
    BndryDv_Outgoing = 2,
    BndryDv_Parallel = 3,
    BndryDv_Invalid = -1
};
#endif
