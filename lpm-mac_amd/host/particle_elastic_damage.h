// Drop-in for the reference header include/particle_elastic_damage.h: the whole API lives in lpm_dropin.hpp.
#pragma once
#include "lpm_dropin.hpp"
