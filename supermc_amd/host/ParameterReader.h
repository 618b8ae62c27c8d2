// ParameterReader -- parameters.dat / command-line reader with the public
// interface of the reference's ParameterReader (ParameterReader.h:22-30):
// "name = value" assignments, text after the comment symbol ignored, names
// trimmed and lower-cased, values read as double.  getVal on a missing name
// prints a message and exits with -1 (ParameterReader.cpp:149-156).
#ifndef SMC_HOST_PARAMETER_READER_H
#define SMC_HOST_PARAMETER_READER_H

#include <string>
#include <vector>

class ParameterReader {
 public:
  ParameterReader();
  ~ParameterReader();
  // one assignment such as " x = 1.1  # comment"
  void phraseOneLine(std::string str, std::string commentSymbol = "#");
  void readFromFile(std::string filename, std::string commentSymbol = "#");
  void readFromArguments(long argc, char* argv[], std::string commentSymbol = "#", long start_from = 1);
  bool exist(std::string name);
  void setVal(std::string name, double value);
  double getVal(std::string name);
  void echo();

 private:
  std::vector<std::string> names_;
  std::vector<double> values_;
  long find(const std::string& name) const;
};

#endif
