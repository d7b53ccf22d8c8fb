#include <KeyParser.h>
#include <console.h>

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace
{
const char* const kBlank = " \t";
const char* const kKeySpace = " \t_!";

void substituteEnvironment(std::string& s)
{
  std::string::size_type at;
  while ((at = s.find("${")) != std::string::npos)
  {
    const auto close = s.find('}', at + 2);
    if (close == std::string::npos) break;
    const std::string name = s.substr(at + 2, close - at - 2);
    const char* value = std::getenv(name.c_str());
    if (value)
      s.replace(at, close - at + 1, value);
    else
    {
      warning("Environment variable '", name,
              " not found. Replaced by empty string.\n"
              "This happened on the following line: ",
              s);
      s.erase(at, close - at + 1);
    }
  }
}

// Text after '=' with surrounding blanks removed; false if none.
bool stringValue(const std::string& line, std::string& out)
{
  const auto eq = line.find('=');
  if (eq == std::string::npos) return false;
  const auto b = line.find_first_not_of(kBlank, eq + 1);
  if (b == std::string::npos) return false;
  const auto e = line.find_last_not_of(kBlank);
  out = line.substr(b, e + 1 - b);
  return true;
}

template<typename T>
bool scalarValue(const std::string& line, T& out)
{
  const auto eq = line.find('=');
  if (eq == std::string::npos) return false;
  std::istringstream in(line.substr(eq + 1));
  T v{};
  in >> v;
  if (in.fail()) return false;
  out = v;
  return true;
}

// {a, b, c} list (or a single bare value).
template<typename T>
bool listValue(const std::string& line, std::vector<T>& out)
{
  const auto eq = line.find('=');
  if (eq == std::string::npos) return false;
  const auto b = line.find_first_not_of(kBlank, eq + 1);
  if (b == std::string::npos) return false;
  std::istringstream in(line.substr(b));
  std::vector<T> values;
  if (line[b] != '{')
  {
    T v{};
    in >> v;
    if (in.fail()) return false;
    out.assign(1, v);
    return true;
  }
  char ch = 0;
  in >> ch; // '{'
  while (true)
  {
    T v{};
    in >> v;
    if (in.fail()) break;
    values.push_back(v);
    in >> std::ws >> ch;
    if (in.fail() || ch != ',') break;
  }
  if (in.fail())
  {
    in.clear();
    in >> std::ws >> ch;
  }
  if (in.fail())
  {
    warning("While reading a vector, expected closing }, but found EOF or "
            "worse.\n",
            "Length of vector returned is ", values.size());
    return false;
  }
  if (ch != '}')
    warning("While reading a vector, expected closing }, but found ", ch,
            " instead.\n", "Length of vector returned is ", values.size());
  out = values;
  return true;
}
}

std::string KeyParser::standardiseKeyword(const std::string& keyword)
{
  const auto first = keyword.find_first_not_of(kKeySpace);
  if (first == std::string::npos) return {};
  const auto last = keyword.find_last_not_of(kKeySpace);
  std::string kw;
  bool prevSpace = false;
  for (auto i = first; i <= last; ++i)
  {
    const char ch = keyword[i];
    if (std::isspace(static_cast<unsigned char>(ch)) || ch == '_' ||
        ch == '!')
    {
      if (!prevSpace) kw += ' ';
      prevSpace = true;
      continue;
    }
    if (ch == '[' && !prevSpace)
      kw += " [";
    else
      kw += static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    prevSpace = false;
  }
  return kw;
}

void KeyParser::add(const std::string& keyword, Kind kind, void* target)
{
  const std::string key = standardiseKeyword(keyword);
  for (auto& e : mEntries)
  {
    if (e.key == key)
    {
      warning("Keyword ", keyword,
              " already registered for parsing. Overwriting previous value.");
      e.kind = kind;
      e.target = target;
      return;
    }
  }
  mEntries.push_back({key, kind, target});
}

void KeyParser::addStartKey(const std::string& k) { add(k, Kind::Start, nullptr); }
void KeyParser::addStopKey(const std::string& k) { add(k, Kind::Stop, nullptr); }
void KeyParser::addKey(const std::string& k, std::string* v) { add(k, Kind::Ascii, v); }
void KeyParser::addKey(const std::string& k, int* v) { add(k, Kind::Int, v); }
void KeyParser::addKey(const std::string& k, float* v) { add(k, Kind::Float, v); }
void KeyParser::addKey(const std::string& k, double* v) { add(k, Kind::Double, v); }
void KeyParser::addKey(const std::string& k, std::vector<int>* v) { add(k, Kind::ListInt, v); }
void KeyParser::addKey(const std::string& k, std::vector<float>* v) { add(k, Kind::ListFloat, v); }
void KeyParser::addKey(const std::string& k, std::vector<double>* v) { add(k, Kind::ListDouble, v); }

const KeyParser::Entry* KeyParser::find(const std::string& key) const
{
  for (const auto& e : mEntries)
    if (e.key == key) return &e;
  return nullptr;
}

const KeyParser::Entry* KeyParser::apply(const std::string& rawLine,
                                         bool writeWarnings)
{
  std::string line = rawLine;
  substituteEnvironment(line);
  const std::string key = standardiseKeyword(line.substr(0, line.find(':')));
  const Entry* e = find(key);
  if (!e)
  {
    if (writeWarnings && !key.empty() && key[0] != ';')
      warning("Unrecognized keyword: ", key);
    return nullptr;
  }
  switch (e->kind)
  {
  case Kind::Start:
    mParsing = true;
    break;
  case Kind::Stop:
    mParsing = false;
    break;
  case Kind::Ascii:
    stringValue(line, *static_cast<std::string*>(e->target));
    break;
  case Kind::Int:
    scalarValue(line, *static_cast<int*>(e->target));
    break;
  case Kind::Float:
    scalarValue(line, *static_cast<float*>(e->target));
    break;
  case Kind::Double:
    scalarValue(line, *static_cast<double*>(e->target));
    break;
  case Kind::ListInt:
    listValue(line, *static_cast<std::vector<int>*>(e->target));
    break;
  case Kind::ListFloat:
    listValue(line, *static_cast<std::vector<float>*>(e->target));
    break;
  case Kind::ListDouble:
    listValue(line, *static_cast<std::vector<double>*>(e->target));
    break;
  }
  return e;
}

bool KeyParser::parse(const std::string& filePath, bool writeWarnings)
{
  std::ifstream in(filePath);
  if (!in) error("Couldn't open file ", filePath);
  mParsing = false;

  // next line with something other than blanks; false at end of input
  auto nextLine = [&](std::string& line) {
    while (true)
    {
      if (!in.good())
      {
        warning("Early EOF or bad file");
        mParsing = false;
        return false;
      }
      std::getline(in, line);
      if (line.find_first_not_of(kBlank) != std::string::npos) return true;
    }
  };

  std::string line;
  if (nextLine(line)) apply(line, writeWarnings);
  if (!mParsing)
  {
    std::string start;
    for (const auto& e : mEntries)
      if (e.kind == Kind::Start) start = e.key;
    if (!start.empty())
      warning("Required first keyword \"", start, "\" not found");
    else
      warning("Data found, but key parser status is \"not parsing\". "
              "Keymap possibly incorrect.");
    return false;
  }
  while (mParsing)
  {
    if (!nextLine(line)) break;
    apply(line, writeWarnings);
  }
  return true;
}
