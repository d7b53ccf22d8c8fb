// console.h -- progress output and error reporting of the FIR host API.
// error() throws std::runtime_error("ERROR: ...\n") and warning() prints
// "WARNING: ..." to stdout, as the reference does (console.inl:64-82); the
// executables catch, print to stderr and return EXIT_FAILURE.
#pragma once

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

inline void echo(const std::string& message)
{
  std::cout << message << std::endl;
}

template<typename... Args>
void print(Args&&... args)
{
  std::ostringstream s;
  (s << ... << std::forward<Args>(args));
  s << '\n';
  std::cout << s.str() << std::flush;
}

template<typename T>
void printValue(const std::string& name, const T& value)
{
  std::cout << name << ": " << value << std::endl;
}

template<typename T>
void printQuotedValue(const std::string& name, const T& value)
{
  std::cout << name << ": \"" << value << "\"" << std::endl;
}

template<typename T>
void printVector(const std::string& name, const std::vector<T>& values)
{
  std::cout << name << ": { ";
  for (size_t i = 0; i < values.size(); ++i)
    std::cout << (i ? ", " : "") << values[i];
  std::cout << " }" << std::endl;
}

inline void printEmptyLine() { std::cout << std::endl; }

template<typename... Args>
void warning(Args&&... args)
{
  std::ostringstream s;
  s << "WARNING: ";
  (s << ... << std::forward<Args>(args));
  s << '\n';
  std::cout << s.str() << std::flush;
}

template<typename... Args>
[[noreturn]] void error(Args&&... args)
{
  std::ostringstream s;
  s << "ERROR: ";
  (s << ... << std::forward<Args>(args));
  s << '\n';
  throw std::runtime_error(s.str());
}

// Wall-clock time of one phase of an executable, printed to stderr as
// "timing: <name> <ms> ms" when the phase ends, if FIR_TIMINGS is set (and
// not "0"). Device work still queued at the end of a phase is counted in
// the phase that waits for it.
class PhaseTimer
{
public:
  explicit PhaseTimer(const char* name)
    : mName(name), mStart(std::chrono::steady_clock::now())
  {
  }
  ~PhaseTimer() { stop(); }
  void stop()
  {
    if (!mName) return;
    if (enabled())
    {
      const double ms = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - mStart)
                          .count();
      std::ostringstream s;
      s << "timing: " << mName << " " << ms << " ms\n";
      std::cerr << s.str() << std::flush;
    }
    mName = nullptr;
  }
  static bool enabled()
  {
    static const bool on = [] {
      const char* e = std::getenv("FIR_TIMINGS");
      return e && *e && std::strcmp(e, "0") != 0;
    }();
    return on;
  }

private:
  const char* mName;
  std::chrono::steady_clock::time_point mStart;
};
