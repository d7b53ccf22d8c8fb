// KeyParser.h -- Interfile-3.3-style "key := value" parser of the FIR host
// API, with the reference's grammar (KeyParser.h:42-157, KeyParser.cc):
//   - keywords are compared after standardisation: case folded, leading
//     and trailing space/tab/'_'/'!' removed, inner runs of white space,
//     '_' and '!' collapsed to one space, " [" inserted before '[';
//   - parsing starts at the start key (which must be the first non-blank
//     line) and ends at the stop key or end of file;
//   - ${NAME} is replaced by the environment variable NAME (removed, with a
//     warning, if unset);
//   - list values are written {a, b, c}; a bare value is a one-element list;
//   - a key without a parsable value leaves its variable unchanged.
#pragma once

#include <string>
#include <vector>

class KeyParser
{
public:
  KeyParser() = default;

  void addStartKey(const std::string& keyword);
  void addStopKey(const std::string& keyword);

  void addKey(const std::string& keyword, std::string* variable);
  void addKey(const std::string& keyword, int* variable);
  void addKey(const std::string& keyword, float* variable);
  void addKey(const std::string& keyword, double* variable);
  void addKey(const std::string& keyword, std::vector<int>* variable);
  void addKey(const std::string& keyword, std::vector<float>* variable);
  void addKey(const std::string& keyword, std::vector<double>* variable);

  // Parse a file. Throws if it cannot be opened; returns false (after a
  // warning) if the first non-blank line is not the start key.
  bool parse(const std::string& filePath, bool writeWarnings = false);

  static std::string standardiseKeyword(const std::string& keyword);

private:
  enum class Kind
  {
    Start,
    Stop,
    Ascii,
    Int,
    Float,
    Double,
    ListInt,
    ListFloat,
    ListDouble
  };

  struct Entry
  {
    std::string key;
    Kind kind;
    void* target;
  };

  void add(const std::string& keyword, Kind kind, void* target);
  const Entry* find(const std::string& standardKey) const;
  // Apply one line; returns the entry kind handled (or nullptr).
  const Entry* apply(const std::string& line, bool writeWarnings);

  std::vector<Entry> mEntries;
  bool mParsing = false;
};
