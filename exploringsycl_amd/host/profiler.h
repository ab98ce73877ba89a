/* Wall-clock profiler used by the host driver (mirrors the interface of
 * TeaLeaf/profiler.h:10-71: named entries with call counts and seconds;
 * not thread safe).  START/STOP_PROFILING are compiled in with
 * -DENABLE_PROFILING, as in the reference. */
#pragma once
#ifndef TEA_PROFILER_H
#define TEA_PROFILER_H

#include <time.h>

#define PROFILER_MAX_NAME 128
#define PROFILER_MAX_ENTRIES 2048

struct ProfileEntry {
  int calls;
  double time;
  char name[PROFILER_MAX_NAME];
};

struct Profile {
  struct timespec profiler_start;
  struct timespec profiler_end;
  int profiler_entry_count;
  struct ProfileEntry profiler_entries[PROFILER_MAX_ENTRIES];
};

void profiler_start_timer(struct Profile* profile);
void profiler_end_timer(struct Profile* profile, const char* entry_name);
void profiler_print_simple_profile(struct Profile* profile);
void profiler_print_full_profile(struct Profile* profile);
int profiler_get_profile_entry(struct Profile* profile, const char* entry_name);

#ifdef ENABLE_PROFILING
#define START_PROFILING(profile) profiler_start_timer(profile)
#define STOP_PROFILING(profile, name) profiler_end_timer(profile, name)
#define PRINT_PROFILING_RESULTS(profile) profiler_print_full_profile(profile)
#else
#define START_PROFILING(profile) ;
#define STOP_PROFILING(profile, name) ;
#define PRINT_PROFILING_RESULTS(profile) ;
#endif

#endif
