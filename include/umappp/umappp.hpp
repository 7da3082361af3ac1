// Drop-in umbrella header for umappp (reference umappp.hpp:9-11), backed by
// libumappp_amd (link with -lumappp_amd).
#ifndef UMAPPP_AMD_UMAPPP_HPP
#define UMAPPP_AMD_UMAPPP_HPP

#include "Options.hpp"
#include "Status.hpp"
#include "initialize.hpp"

#endif
