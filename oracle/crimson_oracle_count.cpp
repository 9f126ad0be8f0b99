// The oracle's counting build (see flopcount.hpp): liboracle_count.so.
#include "flopcount.hpp"

unsigned long long cfo_nflop = 0;

extern "C" {
#include "crimson_oracle.c"

unsigned long long cfo_flop_count(int reset) {
  const unsigned long long n = cfo_nflop;
  if (reset) cfo_nflop = 0;
  return n;
}
}
