// ParameterReader semantics of ParameterReader.cpp:42-169.
#include "ParameterReader.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {

std::string trimmed(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string lowered(std::string s) {
  for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

// first number of the string, as operator>> reads it (arsenal.cpp:425-432)
double first_double(const std::string& s) {
  std::stringstream in(s + " ");
  double v = 0.0;
  in >> v;
  return v;
}

}  // namespace

ParameterReader::ParameterReader() = default;
ParameterReader::~ParameterReader() = default;

long ParameterReader::find(const std::string& name) const {
  const std::string key = lowered(trimmed(name));
  for (size_t i = 0; i < names_.size(); ++i)
    if (names_[i] == key) return (long)i;
  return -1;
}

void ParameterReader::phraseOneLine(std::string str, std::string commentSymbol) {
  if (trimmed(str).empty()) return;
  const std::string eq = str.substr(0, str.find(commentSymbol));
  if (trimmed(eq).empty()) return;
  const size_t pos = eq.find('=');
  if (pos == std::string::npos) {
    std::cout << "ParameterReader::phraseEquationWithoutComments error: \"=\" symbol not found in equation assignment "
              << eq << std::endl;
    std::exit(-1);
  }
  setVal(eq.substr(0, pos), first_double(trimmed(eq.substr(pos + 1))));
}

void ParameterReader::readFromFile(std::string filename, std::string commentSymbol) {
  std::ifstream in(filename.c_str());
  if (!in) {
    std::cout << "ParameterReader::readFromFile error: file " << filename << " does not exist." << std::endl;
    std::exit(-1);
  }
  std::string line;
  while (std::getline(in, line)) phraseOneLine(line, commentSymbol);
}

void ParameterReader::readFromArguments(long argc, char* argv[], std::string commentSymbol, long start_from) {
  for (long i = start_from; i < argc; ++i) phraseOneLine(argv[i], commentSymbol);
}

bool ParameterReader::exist(std::string name) { return find(name) != -1; }

void ParameterReader::setVal(std::string name, double value) {
  const long i = find(name);
  if (i == -1) {
    names_.push_back(lowered(trimmed(name)));
    values_.push_back(value);
  } else {
    values_[i] = value;
  }
}

double ParameterReader::getVal(std::string name) {
  const long i = find(name);
  if (i != -1) return values_[i];
  std::cout << "ParameterReader::getVal error: parameter with name " << name << " not found." << std::endl;
  std::exit(-1);
}

void ParameterReader::echo() {
  if (names_.empty()) return;
  for (size_t i = 0; i < names_.size(); ++i) std::cout << names_[i] << "=" << values_[i] << "  ";
  std::cout << std::endl;
}
