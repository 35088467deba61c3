/* Wavelet/WAVELET_1D.h of doseongju/C_pp-library: WAVELET_1D is defined in awcm_api.h. */
#pragma once
#include "awcm_api.h"
