// Logging and fatal errors (behaviour of TeaLeaf/shared.c:4-111): messages
// go to stdout and to the tea.out log on rank 0; die() reports the location
// and ends the run through abort_comms().
#include "shared.h"

#include "comms.h"
#include "settings.h"
#include "tea_host.h"

void initialise_log(struct Settings* settings) {
  if (settings->rank != MASTER) return;
  if (!g_opts.quiet) printf("Opening %s as log file.\n", settings->tea_out_filename);
  settings->tea_out_fp = fopen(settings->tea_out_filename, "w");
  if (!settings->tea_out_fp)
    die(__LINE__, __FILE__, "Could not open log %s\n", settings->tea_out_filename);
}

void print_and_log(struct Settings* settings, const char* format, ...) {
  if (settings->rank != MASTER) return;
  if (!g_opts.quiet) {
    va_list args;
    va_start(args, format);
    vprintf(format, args);
    va_end(args);
  }
  if (!settings->tea_out_fp)
    die(__LINE__, __FILE__, "Attempted to write to log before it was initialised\n");
  va_list args2;
  va_start(args2, format);
  vfprintf(settings->tea_out_fp, format, args2);
  va_end(args2);
}

void print_to_log(struct Settings* settings, const char* format, ...) {
  if (settings->rank != MASTER) return;
  if (!settings->tea_out_fp)
    die(__LINE__, __FILE__, "Attempted to write to log before it was initialised\n");
  va_list args;
  va_start(args, format);
  vfprintf(settings->tea_out_fp, format, args);
  va_end(args);
}

void plot_2d(int x, int y, double* buffer, const char* name) {
  FILE* fp = fopen("plot2d.dat", "wb");
  if (!fp) {
    printf("Could not open plot file.\n");
    return;
  }
  double sum = 0.0;
  for (int jj = 0; jj < y; ++jj)
    for (int kk = 0; kk < x; ++kk) {
      const double v = buffer[kk + jj * x];
      fprintf(fp, "%d %d %.12E\n", kk, jj, v);
      sum += v;
    }
  printf("%s: %.12E\n", name, sum);
  fclose(fp);
}

void die(int lineNum, const char* file, const char* format, ...) {
  fflush(stdout);
  fprintf(stderr, "\nError at line %d in %s:\n", lineNum, file);
  va_list args;
  va_start(args, format);
  vfprintf(stderr, format, args);
  va_end(args);
  fflush(stderr);
  abort_comms();
}
