// operations.h -- image operations of the FIR host API (reference
// operations.h:7-22). convolve, cutCircle and HounsfieldToMuMap run on the
// GPU; applyMask is host-side preprocessing.
#pragma once

#include <VolData.h>

#include <vector>

namespace operations
{
// Separable Gaussian (FWHM in mm per axis) on every frame, with border
// renormalisation; with cutRadius > 0, voxels at radius >=
// cutRadius - 5 * max(FWHM_x, FWHM_y) keep their original value. Skipped
// unless all three FWHM are > 0. Leaves the last frame active.
void convolve(VolData& vol, std::vector<float> fwhmXYZ, float cutRadius = 0);

// Zero the voxels of the active frame whose centre is farther than
// cutRadius from the image axis; no-op for cutRadius <= 0.
void cutCircle(VolData& vol, float cutRadius = 0);

// Zero the voxels of the active frame where the mask is <= 0.
void applyMask(VolData& vol, VolData& maskVol);

// Hounsfield units to linear attenuation (1/mm) on the active frame.
void HounsfieldToMuMap(VolData& vol);
}
