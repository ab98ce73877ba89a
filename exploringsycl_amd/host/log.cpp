// The host program's log and its fatal-error exit.
//
// Rank 0 formats each message once, appends it to the tea.out log (opened on
// first use, so nothing has to be set up before the first message) and, for
// print_and_log and unless --quiet was given, echoes it to stdout.  Other
// ranks stay silent.  die() reports where it was called from and ends the
// run through the communication layer.
#include <stdarg.h>

#include <string>

#include "comms.h"
#include "settings.h"
#include "shared.h"
#include "tea_host.h"

namespace {

FILE* log_file(Settings* settings) {
  if (settings->tea_out_fp == nullptr) {
    settings->tea_out_fp = fopen(settings->tea_out_filename, "w");
    if (settings->tea_out_fp == nullptr)
      die(__LINE__, __FILE__, "cannot open the log file '%s' for writing\n",
          settings->tea_out_filename);
  }
  return settings->tea_out_fp;
}

// printf-style formatting into a std::string.
std::string vformat(const char* format, va_list args) {
  va_list probe;
  va_copy(probe, args);
  const int n = vsnprintf(nullptr, 0, format, probe);
  va_end(probe);
  if (n <= 0) return std::string();
  std::string text(static_cast<size_t>(n) + 1, '\0');
  vsnprintf(&text[0], text.size(), format, args);
  text.resize(static_cast<size_t>(n));
  return text;
}

void emit(Settings* settings, bool echo, const char* format, va_list args) {
  if (settings->rank != MASTER) return;
  const std::string text = vformat(format, args);
  fwrite(text.data(), 1, text.size(), log_file(settings));
  if (echo && !g_opts.quiet) fwrite(text.data(), 1, text.size(), stdout);
}

}  // namespace

void initialise_log(Settings* settings) {
  if (settings->rank == MASTER) log_file(settings);
}

void print_and_log(Settings* settings, const char* format, ...) {
  va_list args;
  va_start(args, format);
  emit(settings, true, format, args);
  va_end(args);
}

void print_to_log(Settings* settings, const char* format, ...) {
  va_list args;
  va_start(args, format);
  emit(settings, false, format, args);
  va_end(args);
}

void die(int line, const char* file, const char* format, ...) {
  va_list args;
  va_start(args, format);
  const std::string why = vformat(format, args);
  va_end(args);
  fflush(stdout);
  fprintf(stderr, "tealeaf: fatal (%s:%d): %s", file, line, why.c_str());
  if (why.empty() || why.back() != '\n') fputc('\n', stderr);
  fflush(stderr);
  abort_comms();
}
