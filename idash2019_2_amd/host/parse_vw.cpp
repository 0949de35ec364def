#include "parse_vw.h"

#include <cstdio>
#include <cstdlib>

std::unordered_map<std::string, int32_t> read(const std::string &file_name) {
  FILE *f = std::fopen(file_name.c_str(), "r");
  if (!f) {
    std::fprintf(stderr, "Cannot open file '%s'\n", file_name.c_str());
    std::abort();
  }
  std::unordered_map<std::string, int32_t> coefs;
  char line[256], name[256];
  int32_t value;
  while (std::fgets(line, sizeof(line), f))
    if (std::sscanf(line, "%255s %d", name, &value) == 2) coefs[name] = value;
  std::fclose(f);
  return coefs;
}
