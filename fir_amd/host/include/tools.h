// tools.h -- string, endianness and path helpers of the FIR host API
// (reference tools.h:1-33).
#pragma once

#include <climits>
#include <cstddef>
#include <string>

std::string strToUpper(const std::string& in);

bool systemIsLittleEndian();

template<typename T>
T swapEndianness(T u)
{
  static_assert(CHAR_BIT == 8, "CHAR_BIT != 8");
  unsigned char b[sizeof(T)];
  __builtin_memcpy(b, &u, sizeof(T));
  for (size_t k = 0; k < sizeof(T) / 2; ++k)
  {
    const unsigned char t = b[k];
    b[k] = b[sizeof(T) - 1 - k];
    b[sizeof(T) - 1 - k] = t;
  }
  T out;
  __builtin_memcpy(&out, b, sizeof(T));
  return out;
}

// A relative data-file path is taken relative to the header's directory
// (changed in place).
void addPath(const std::string& headerFile, std::string& dataFile);

// Host threads are not used by the projectors (they run on the GPU); these
// report 1 and 0 so that callers printing them keep working.
int getNThreads();
int getCurrentThread();
