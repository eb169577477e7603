// putil.cpp -- numeric command-line arguments with magnitude suffixes.
//
// Same accepted syntax and status codes as the reference's parse_uint64
// (putil.c:34-116): strtoull base 0, then an optional suffix
//   K M G T P  -> x 10^3 .. 10^15      e<n> / E<n> -> x 10^n
//   k m g t p  -> x 2^10 .. 2^50       b<n> / B<n> -> x 2^n
// Returns 0 on success, -1 if the text cannot be parsed, -2 if out of range.
#include <errno.h>
#include <stdint.h>
#include <stdlib.h>

#include "putil.h"

static int scale(uint64_t* num, uint64_t hi, uint64_t factor, unsigned times) {
  while (times--) {
    if (*num > hi / factor) return -2;
    *num *= factor;
  }
  return 0;
}

extern "C" int clw_parse_uint64(uint64_t* result, const char* str, uint64_t lo, uint64_t hi) {
  char* tail = nullptr;
  errno = 0;
  uint64_t num = strtoull(str, &tail, 0);
  if (errno != 0 || num > hi) return -2;
  const char c = *tail;
  int rc = 0;
  if (c == '\0') {
    // plain number
  } else if (c == 'K' || c == 'M' || c == 'G' || c == 'T' || c == 'P') {
    if (tail[1] != '\0') return -1;
    const unsigned steps = c == 'K' ? 1 : c == 'M' ? 2 : c == 'G' ? 3 : c == 'T' ? 4 : 5;
    rc = scale(&num, hi, 1000, steps);
  } else if (c == 'k' || c == 'm' || c == 'g' || c == 't' || c == 'p') {
    if (tail[1] != '\0') return -1;
    const unsigned sh = c == 'k' ? 10 : c == 'm' ? 20 : c == 'g' ? 30 : c == 't' ? 40 : 50;
    if (num > (hi >> sh)) return -2;
    num <<= sh;
  } else if (c == 'e' || c == 'E' || c == 'b' || c == 'B') {
    errno = 0;
    const unsigned long ex = strtoul(tail + 1, &tail, 0);
    if (errno != 0) return -2;
    if (*tail != '\0') return -1;
    rc = (c == 'e' || c == 'E') ? scale(&num, hi, 10, (unsigned)ex) : scale(&num, hi, 2, (unsigned)ex);
  } else {
    return -1;
  }
  if (rc) return rc;
  if (num < lo) return -2;
  *result = num;
  return 0;
}
