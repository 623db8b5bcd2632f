// init.hpp -- host-side initial flow fields (clVariablesLB::iniDists(Umax)).
#pragma once
#include "../../include/kofm.h"
#include "geometry.hpp"

namespace kofm {

// rho, ux, uy [D] of clVariablesLB::iniDists(Umax) (clVariablesLB.cpp:427-496)
// on geometry g; nodes it does not visit are left as they are.
void iniPoiseuille(const Geometry& g, const kofm_ini_flow& p, double* rho, double* ux, double* uy);

}  // namespace kofm
