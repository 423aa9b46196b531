// SmoothWindow.h -- apodisation windows of the ramp filter.  The reference's
// Include/SmoothWindow.h is an empty stub; these are the windows this build
// defines (frequency response |f| W(f), f_N = 1 / (2 ds)):
//   NONE         W = 1                      (reference Ram-Lak taps)
//   HANN         W = 0.5 (1 + cos(pi f/f_N))
//   SHEPP_LOGAN  W = sinc(f / (2 f_N))
#pragma once

#include "lct.h"

enum class SmoothWindow { NONE = LCT_WINDOW_NONE, HANN = LCT_WINDOW_HANN, SHEPP_LOGAN = LCT_WINDOW_SHEPP_LOGAN };
