// Fused step driver.  (Placeholder body: the phase-by-phase sequence.)
#include "ark_internal.hpp"

ark_status ark_step_fused(ark_ctx* c, int nsteps) { return ark_step_phases(c, nsteps); }
