// projections.h -- projection operators of the FIR host API, same
// signatures as the reference (projections.h:9-24), computed on the GPU.
#pragma once

#include <ProjData.h>
#include <ScannerData.h>
#include <VolData.h>

namespace projections
{
// Forward projection of the active frame of inputVol into every bin.
void forward(const VolData& inputVol, const ScannerData& scanner,
             ProjData& outputProj);

// Backprojection; outputVol must have nSubsets frames: frame s receives
// the LORs of subset s (views v with v % nSubsets == s).
void backward(const ProjData& inputProj, const ScannerData& scanner,
              VolData& outputVol, int nSubsets = 1);

// Backprojection of ones, one frame per subset.
void computeSensitivityVol(const ProjData& proj, const ScannerData& scanner,
                           VolData& initializedSensVol, int nSubsets = 1);
}
