// Launch-shape autotuner.
#pragma once
#include <functional>
#include <string>

#include "common.hpp"

namespace binder {

// Per (operation, problem key): the fastest of a few candidate launch
// shapes, timed with HIP events on the caller's stream the first time the
// operation runs on that problem while tuning is enabled, then cached for
// the process.  Callers only offer candidates that leave every result bit for
// bit unchanged (voxel ranges whose partial sums are exact integers, tiles
// computed independently of the block that runs them), so a tuned run
// computes what the default shape computes.
class Autotuner {
 public:
  static Autotuner &get();
  bool enabled() const;
  // 0 off, 1 tune by timing, 2 take the last non-default candidate untimed
  // (checks that a non-default shape gives the default's results)
  void set_mode(int mode);
  // The cached choice for (op, key); else, when tuning is enabled and `s`
  // is not being captured, each candidate is run (one warm-up, then the
  // best of three timed runs) through run(candidate) and the fastest is
  // cached and returned; otherwise `dflt`.
  int choose(const char *op, long long key, const int *cands, int ncands, int dflt,
             hipStream_t s, const std::function<void(int)> &run);
  // One line per tuned problem: op key choice default ms(choice) ms(default).
  std::string report() const;
  void clear();
};

}  // namespace binder
