// model_host_device.h -- macroscopic variable kinds with the values of the
// reference (Src/model_host_device.h:6-19).
#ifndef MODEL_HOST_DEVICE_H
#define MODEL_HOST_DEVICE_H
enum VariableTypes {
    Variable_Rho = 0,
    Variable_U = 1,
    Variable_V = 2,
    Variable_W = 3,
    Variable_T = 4,
    Variable_Qx = 5,
    Variable_Qy = 6,
    Variable_Qz = 7,
    Variable_U_Force = 8,
    Variable_V_Force = 9,
    Variable_W_Force = 10,
};
#endif
