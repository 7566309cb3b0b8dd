// timer.h -- drop-in of lib/timer.h: a wall-clock timer (microsecond
// gettimeofday stamps) and the global total_timer started at static
// initialisation.
#ifndef DS_DROPIN_TIMER_H
#define DS_DROPIN_TIMER_H

#include <sys/time.h>

#include <cstddef>

typedef struct timeval tv;

class Timer {
 public:
  Timer();
  static tv get_time();
  // seconds from t0 to t1
  static double interval(const tv t0, const tv t1);
  void reset();
  double elapsed() const;

 private:
  tv t0;
};

extern Timer total_timer;

#endif
