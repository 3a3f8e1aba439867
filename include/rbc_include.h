/* Umbrella header, the counterpart of the reference's rbc_include.h:1-13. */
#ifndef RBC_INCLUDE_H
#define RBC_INCLUDE_H

#include "rbc_defs.h"
#include "rbc_hip.h"
#include "rbc_utils.h"
#include "rbc.h"
#include "brute.h"

#endif
