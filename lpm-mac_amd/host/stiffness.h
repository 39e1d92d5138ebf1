// Drop-in for the reference header include/stiffness.h: the whole API lives in lpm_dropin.hpp.
#pragma once
#include "lpm_dropin.hpp"
