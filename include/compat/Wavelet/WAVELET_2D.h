/* Wavelet/WAVELET_2D.h of doseongju/C_pp-library: WAVELET_2D is defined in awcm_api.h. */
#pragma once
#include "awcm_api.h"
