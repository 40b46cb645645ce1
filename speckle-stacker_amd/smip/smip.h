// smip.h -- the drop-in C++ surface of the MI355X speckle-masking engine in one
// include: Array2, Bispectrum, reconstruct_phases / calc_phase / NextRecoIndex,
// CrossCorrelation and the window functions, in namespace smip.
#pragma once

#include "array2.h"
#include "bispectrum.h"
#include "crosscorrel.h"
#include "phasemap.h"
#include "phasereco.h"
#include "window_function.h"
