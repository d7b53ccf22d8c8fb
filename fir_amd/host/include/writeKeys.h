// writeKeys.h -- "key := value" writers of the FIR host API. Doubles are
// written with "%.7f" (reference writeKeys.cc:12-21), everything else with
// operator<<.
#pragma once

#include <cstdio>
#include <ostream>
#include <string>

inline void writeKey(std::ostream& os, const std::string& key)
{
  os << key << " :=" << '\n';
}

inline void writeKey(std::ostream& os, const std::string& key, double value)
{
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.7f", value);
  os << key << " := " << buf << '\n';
}

template<typename T>
void writeKey(std::ostream& os, const std::string& key, const T& value)
{
  os << key << " := " << value << '\n';
}
