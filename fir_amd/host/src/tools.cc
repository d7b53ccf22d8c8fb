#include <tools.h>

#include <cctype>
#include <cstdint>

std::string strToUpper(const std::string& in)
{
  std::string out(in);
  for (auto& ch : out)
    ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
  return out;
}

bool systemIsLittleEndian()
{
  const std::uint16_t probe = 1;
  unsigned char first;
  __builtin_memcpy(&first, &probe, 1);
  return first == 1;
}

void addPath(const std::string& headerFile, std::string& dataFile)
{
  if (dataFile.empty() || dataFile[0] == '/') return;
  const auto slash = headerFile.rfind('/');
  if (slash != std::string::npos)
    dataFile = headerFile.substr(0, slash + 1) + dataFile;
}

int getNThreads() { return 1; }

int getCurrentThread() { return 0; }
