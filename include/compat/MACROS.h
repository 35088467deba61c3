/* MACROS.h of doseongju/C_pp-library: its classes are defined in awcm_api.h. */
#pragma once
#include "awcm_api.h"
