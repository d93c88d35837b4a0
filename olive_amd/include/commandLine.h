// Command-line parsing of the Olive C++ API (commandLine.h:42-101 of the
// reference): positional arguments and `-option value` pairs.
#ifndef OLIVE_COMMAND_LINE_H
#define OLIVE_COMMAND_LINE_H

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

class CommandLine {
 public:
  CommandLine(int c, char **v, std::string s) : argc_(c), argv_(v), example_(s) {}

  void badArgument() {
    std::cout << "usage: " << argv_[0] << " " << example_ << std::endl;
    abort();
  }

  /** Positional argument i (0 is the first after the program name). */
  char *getArgument(int i) {
    if (argc_ < i + 2) badArgument();
    return argv_[i + 1];
  }

  bool getOption(std::string option) {
    for (int i = 1; i < argc_; i++)
      if ((std::string)argv_[i] == option) return true;
    return false;
  }

  int getOptionIntValue(std::string option, int defaultValue) {
    for (int i = 1; i < argc_ - 1; i++)
      if ((std::string)argv_[i] == option) return atoi(argv_[i + 1]);
    return defaultValue;
  }

  long getOptionLongValue(std::string option, long defaultValue) {
    for (int i = 1; i < argc_ - 1; i++)
      if ((std::string)argv_[i] == option) return atol(argv_[i + 1]);
    return defaultValue;
  }

  const char *getOptionStringValue(std::string option, const char *defaultValue) {
    for (int i = 1; i < argc_ - 1; i++)
      if ((std::string)argv_[i] == option) return argv_[i + 1];
    return defaultValue;
  }

  double getOptionDoubleValue(std::string option, double defaultValue) {
    for (int i = 1; i < argc_ - 1; i++)
      if ((std::string)argv_[i] == option) {
        double val;
        if (sscanf(argv_[i + 1], "%lf", &val) == EOF) badArgument();
        return val;
      }
    return defaultValue;
  }

 private:
  int argc_;
  char **argv_;
  std::string example_;
};

#endif  // OLIVE_COMMAND_LINE_H
