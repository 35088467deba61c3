/* Problems/1D/PROBLEM_SHOCKTUBE.h of doseongju/C_pp-library: SHOCKTUBE is defined
 * in awcm_api.h; this header declares the problem's BOUNDARY_MANAGER, as the
 * reference's does. */
#pragma once
#include "awcm_api.h"

class BOUNDARY_MANAGER : public awcm_api_detail::Extrap1D {
public:
    BOUNDARY_MANAGER(FIELD_UNIFORM_1D<VECTOR_3D<double>>& Q_) : Extrap1D(Q_) {}
};
