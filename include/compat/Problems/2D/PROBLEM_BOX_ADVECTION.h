/* Problems/2D/PROBLEM_BOX_ADVECTION.h of doseongju/C_pp-library: BOX_ADVECTION is
 * defined in awcm_api.h; this header declares the problem's (periodic, scalar)
 * BOUNDARY_MANAGER, as the reference's does. */
#pragma once
#include "awcm_api.h"

class BOUNDARY_MANAGER : public awcm_api_detail::Periodic2D {
public:
    BOUNDARY_MANAGER(FIELD_UNIFORM_2D<double>& Q_) : Periodic2D(Q_) {}
};
