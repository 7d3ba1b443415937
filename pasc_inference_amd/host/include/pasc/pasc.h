// Umbrella header of the C++ host layer over libpasc_hip.so.
#pragma once

#include "common.h"
#include "decomposition.h"
#include "imagedata.h"
#include "model.h"
#include "solvers.h"
#include "tsdata.h"
#include "tssolver.h"
