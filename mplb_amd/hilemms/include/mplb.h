// mplb.h -- umbrella header of the HiLeMMS API (reference Src/mplb.h:33-47)
#ifndef MPLB_H
#define MPLB_H
#include "boundary.h"
#include "configuration.h"
#include "evolution.h"
#include "flowfield.h"
#include "model.h"
#include "scheme.h"
#include "type.h"
#endif
