// boinc_stub.h -- the subset of the BOINC client API that the Wilson-prime
// host calls (main.cpp:20-330 and cl_wilson.cpp:50-80, 290-320, 356-470 of the
// reference), with standalone behaviour: no client, files by their own names,
// progress on stderr, a checkpoint every 60 s. A volunteer-computing build
// links the real BOINC library in place of boinc_stub.cpp; the call sites do
// not change.
#pragma once
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

int boinc_init(void);
int boinc_finish(int status);  // never returns
int boinc_is_standalone(void);
int boinc_resolve_filename(const char* virtual_name, char* physical_name, int len);
FILE* boinc_fopen(const char* path, const char* mode);
int boinc_fraction_done(double fraction);
double boinc_get_fraction_done(void);
double boinc_elapsed_time(void);
int boinc_time_to_checkpoint(void);
int boinc_checkpoint_completed(void);
int boinc_begin_critical_section(void);
int boinc_end_critical_section(void);
int boinc_send_trickle_up(char* variety, char* text);

#ifdef __cplusplus
}
#endif
