// solvers.h -- internal include of the public solver declarations.
#pragma once
#include "../../include/pd_solvers.h"
