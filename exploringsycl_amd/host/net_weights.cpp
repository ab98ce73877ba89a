// Reader of the learned preconditioner's network file (format in tea_host.h).
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "tea_host.h"

bool read_net_weights(const char* path, NetWeights* out, std::string* why) {
  FILE* fp = fopen(path, "rb");
  if (!fp) {
    *why = std::string("cannot open ") + path;
    return false;
  }
  char line[512];
  if (!fgets(line, sizeof line, fp)) {
    fclose(fp);
    *why = "empty file";
    return false;
  }
  int version = 0, count = 0, coarsen = 0;
  float nm[4];
  const int got = sscanf(line, "tea-nn-weights %d %d %d %f %f %f %f", &version,
                         &count, &coarsen, &nm[0], &nm[1], &nm[2], &nm[3]);
  if (got != 7 || version != 1 || count <= 0 || coarsen <= 0 || !(nm[1] > 0.f) ||
      !(nm[3] > 0.f)) {
    fclose(fp);
    *why = "malformed header line";
    return false;
  }
  out->params.assign((size_t)count, 0.f);
  const size_t n = fread(out->params.data(), sizeof(float), (size_t)count, fp);
  fclose(fp);
  if (n != (size_t)count) {
    *why = "file holds fewer parameters than its header states";
    return false;
  }
  for (float v : out->params)
    if (!std::isfinite(v)) {
      *why = "non-finite parameter";
      return false;
    }
  out->coarsen = coarsen;
  memcpy(out->norm, nm, sizeof nm);
  return true;
}
