// boinc_stub.cpp -- standalone implementation of the BOINC API subset in
// boinc_stub.h: the program runs without a BOINC client.
#include "boinc_stub.h"

#include <stdlib.h>
#include <string.h>

#include <chrono>

namespace {
using clock_type = std::chrono::steady_clock;
clock_type::time_point g_start = clock_type::now();
clock_type::time_point g_last_checkpoint = clock_type::now();
double g_fraction = 0.0;
const double kCheckpointSeconds = 60.0;
}  // namespace

extern "C" {

int boinc_init(void) {
  g_start = g_last_checkpoint = clock_type::now();
  g_fraction = 0.0;
  return 0;
}

int boinc_finish(int status) {
  fflush(stdout);
  fflush(stderr);
  exit(status);
}

int boinc_is_standalone(void) { return 1; }

int boinc_resolve_filename(const char* virtual_name, char* physical_name, int len) {
  if (!virtual_name || !physical_name || len <= 0) return -1;
  strncpy(physical_name, virtual_name, (size_t)len - 1);
  physical_name[len - 1] = 0;
  return 0;
}

FILE* boinc_fopen(const char* path, const char* mode) { return fopen(path, mode); }

int boinc_fraction_done(double fraction) {
  g_fraction = fraction < 0 ? 0 : (fraction > 1 ? 1 : fraction);
  return 0;
}

double boinc_get_fraction_done(void) { return g_fraction; }

double boinc_elapsed_time(void) {
  return std::chrono::duration<double>(clock_type::now() - g_start).count();
}

int boinc_time_to_checkpoint(void) {
  return std::chrono::duration<double>(clock_type::now() - g_last_checkpoint).count() >= kCheckpointSeconds;
}

int boinc_checkpoint_completed(void) {
  g_last_checkpoint = clock_type::now();
  return 0;
}

int boinc_begin_critical_section(void) { return 0; }
int boinc_end_critical_section(void) { return 0; }

// Standalone: there is no server to send progress to.
int boinc_send_trickle_up(char*, char*) { return 0; }

}  // extern "C"
