// Logging of the Olive C++ API (logging.h:42-101 of the reference):
// LOG(level) << ... writes one line to stderr when the level is at or below
// the reporting level.  Not thread-safe.
#ifndef OLIVE_LOGGING_H
#define OLIVE_LOGGING_H

#include <sstream>
#include <string>

#include "common.h"

enum LogLevel { ERROR = 0, WARNING, INFO, DEBUG, DEBUG1, DEBUG2, DEBUG3 };

class Logging {
 public:
  ~Logging() {
    if (messageLevel_ <= ReportingLevel()) {
      os_ << std::endl;
      fprintf(stderr, "%s", os_.str().c_str());
      fflush(stderr);
    }
  }
  std::ostringstream &Get(LogLevel level = INFO) {
    os_ << name(level) << ": ";
    messageLevel_ = level;
    return os_;
  }
  static LogLevel &ReportingLevel() {
    static LogLevel level = DEBUG;
    return level;
  }

 private:
  static const char *name(LogLevel l) {
    switch (l) {
      case ERROR: return "ERROR";
      case WARNING: return "WARNING";
      case INFO: return "INFO";
      case DEBUG: return "DEBUG";
      case DEBUG1: return "DEBUG1";
      case DEBUG2: return "DEBUG2";
      case DEBUG3: return "DEBUG3";
    }
    return "UNKNOWN";
  }
  std::ostringstream os_;
  LogLevel messageLevel_ = INFO;
};

#define LOG(level) Logging().Get(level)

#endif  // OLIVE_LOGGING_H
