// Named wall-clock timers (interface of TeaLeaf/profiler.c:10-118).
#include "profiler.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double elapsed_seconds(const struct timespec& a, const struct timespec& b) {
  return (double)(b.tv_sec - a.tv_sec) + (double)(b.tv_nsec - a.tv_nsec) * 1e-9;
}

void profiler_start_timer(struct Profile* profile) {
  clock_gettime(CLOCK_MONOTONIC, &profile->profiler_start);
}

void profiler_end_timer(struct Profile* profile, const char* entry_name) {
  clock_gettime(CLOCK_MONOTONIC, &profile->profiler_end);
  int ii = 0;
  while (ii < profile->profiler_entry_count &&
         strcmp(profile->profiler_entries[ii].name, entry_name) != 0)
    ++ii;
  if (ii >= PROFILER_MAX_ENTRIES) {
    printf("Attempted to profile too many entries, maximum is %d\n",
           PROFILER_MAX_ENTRIES);
    exit(1);
  }
  struct ProfileEntry* e = &profile->profiler_entries[ii];
  if (ii == profile->profiler_entry_count) {
    profile->profiler_entry_count++;
    snprintf(e->name, PROFILER_MAX_NAME, "%s", entry_name);
    e->time = 0.0;
    e->calls = 0;
  }
  e->time += elapsed_seconds(profile->profiler_start, profile->profiler_end);
  e->calls++;
}

void profiler_print_full_profile(struct Profile* profile) {
  printf("\n-------------------------------------------------------------\n");
  printf("\nProfiling Results:\n\n");
  printf("%-30s%8s%20s\n", "Kernel Name", "Calls", "Runtime (s)");
  double total = 0.0;
  for (int ii = 0; ii < profile->profiler_entry_count; ++ii) {
    const struct ProfileEntry* e = &profile->profiler_entries[ii];
    total += e->time;
    printf("%-30s%8d%20.03f\n", e->name, e->calls, e->time);
  }
  printf("\nTotal elapsed time: %.03fs, entries * are excluded.\n", total);
  printf("\n-------------------------------------------------------------\n\n");
}

void profiler_print_simple_profile(struct Profile* profile) {
  for (int ii = 0; ii < profile->profiler_entry_count; ++ii)
    printf("%s: %.3lfs (%d calls)\n", profile->profiler_entries[ii].name,
           profile->profiler_entries[ii].time,
           profile->profiler_entries[ii].calls);
}

int profiler_get_profile_entry(struct Profile* profile, const char* entry_name) {
  for (int ii = 0; ii < profile->profiler_entry_count; ++ii)
    if (strcmp(profile->profiler_entries[ii].name, entry_name) == 0) return ii;
  printf("Attempted to retrieve missing profile entry %s\n", entry_name);
  exit(1);
}
