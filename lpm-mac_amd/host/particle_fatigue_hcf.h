// Drop-in for the reference header include/particle_fatigue_hcf.h: the whole API lives in lpm_dropin.hpp.
#pragma once
#include "lpm_dropin.hpp"
