// scheme.h -- scheme selection and the stencil names Apps pass to
// ops_arg_dat (reference Src/scheme.h:40-89).
#ifndef SCHEME_H
#define SCHEME_H
#include "flowfield.h"
#include "model.h"
#include "type.h"

extern ops_stencil LOCALSTENCIL;
extern ops_stencil ONEPTREGULARSTENCIL;
extern ops_stencil TWOPTREGULARSTENCIL;
extern ops_stencil ONEPTLATTICESTENCIL;

enum SchemeType {
    Scheme_E1st2nd = 1,
    Scheme_I1st2nd = -1,
    Scheme_StreamCollision = 10,
    Scheme_StreamCollision_Swap = 11,
};
void SetupCommonStencils();
void DefineScheme(const SchemeType scheme);
int SchemeHaloNum();
void SetSchemeHaloNum(const int schemeHaloNum);
SchemeType Scheme();
#ifdef OPS_3D
void PredefinedStream3D();
#endif
#ifdef OPS_2D
void Stream();
#endif
#endif
