// macros.h -- numeric constants and loop helpers of the FIR host API
// (reference macros.h:1-48).
#pragma once

#define ABS(a) ((a) < 0 ? (-(a)) : (a))
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

// Clamp threshold of the element-wise operations and the walk.
#define EPSILON 0.00001

#define PI 3.141592653589793

#define LOOP(INDEX, FIRST, LAST) \
  for (int INDEX = (FIRST); INDEX <= (LAST); ++INDEX)

#define LOOP_SEG(SEG, PROJ) \
  LOOP(SEG, -(PROJ).getGeometry().segOffset, (PROJ).getGeometry().segOffset)

#define LOOP_VIEW(VIEW, PROJ) LOOP(VIEW, 0, (PROJ).getGeometry().nViews - 1)

#define LOOP_AXIAL(AXIAL, PROJ, SEG) \
  LOOP(AXIAL, 0, (PROJ).getGeometry().getNAxialCoords(SEG) - 1)

#define LOOP_TANG(TANG, PROJ)                        \
  LOOP(TANG, -(PROJ).getGeometry().tangCoordOffset, \
       -(PROJ).getGeometry().tangCoordOffset +      \
         (PROJ).getHeader().nTangCoords - 1)

// Bin of segment SEG at index BIN_INDEX within the segment.
#define BIN(PROJ, SEG, BIN_INDEX) \
  (PROJ).getDataArray()[(SEG) + (PROJ).getGeometry().segOffset][BIN_INDEX]
