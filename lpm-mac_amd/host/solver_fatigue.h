// Drop-in for the reference header include/solver_fatigue.h: the whole API lives in lpm_dropin.hpp.
#pragma once
#include "lpm_dropin.hpp"
