// Wall-clock helpers of the Olive C++ API (timer.h:41-74 of the reference).
#ifndef OLIVE_TIMER_H
#define OLIVE_TIMER_H

#include <sys/time.h>

/** Current time in milliseconds. */
inline double getTimeMillis() {
  timeval t;
  gettimeofday(&t, NULL);
  return t.tv_sec * 1000.0 + t.tv_usec / 1000.0;
}

/** Current time in seconds. */
inline double getTimeSeconds() {
  timeval t;
  gettimeofday(&t, NULL);
  return t.tv_sec + t.tv_usec / 1000000.0;
}

/** Elapsed time since the previous call (or start()). */
class Stopwatch {
 public:
  void start() { timer_ = getTimeMillis(); }
  double getElapsedMillis() {
    double now = getTimeMillis();
    double elapsed = now - timer_;
    timer_ = now;
    return elapsed;
  }

 private:
  double timer_ = 0.0;
};

#endif  // OLIVE_TIMER_H
