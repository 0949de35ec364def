// Model file reader (the reference's eval/parse_vw.h:1-14).
#ifndef IHE_HOST_PARSE_VW_H
#define IHE_HOST_PARSE_VW_H

#include <cstdint>
#include <string>
#include <unordered_map>

// Reads a .hr model file: one "<name> <integer>" pair per line, name being
// "<pos>_<snp>" or "Constant"; the integer is read as %d, so "-123.0" gives
// -123 (eval/parse_vw.cpp:7-29). Lines without a name and a number are
// skipped. A missing file prints to stderr and aborts.
std::unordered_map<std::string, int32_t> read(const std::string &file_name);

#endif  // IHE_HOST_PARSE_VW_H
