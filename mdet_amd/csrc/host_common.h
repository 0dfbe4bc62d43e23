// Shared host-side definitions of the mdet TNF library: error reporting and
// the canonical tetranucleotide column tables.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mdet/tnf.h"

namespace mdet {

// Thread-local error message behind mdet_last_error().
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// 2-bit base code used everywhere in this library: A=0 C=1 G=2 T=3, so the
// complement of b is 3-b.  A 4-mer code puts the FIRST base in the LOW bits:
// code = b0 | b1<<2 | b2<<4 | b3<<6 (the order bases sit in a packed word).
// column_of_code[code] is the TNF column the 4-mer counts into.
struct ColumnTables {
    int16_t column_of_code[256];
    bool ready;
};
const ColumnTables& column_tables();

}  // namespace mdet

#define MDET_FAIL(code, ...)            \
    do {                                \
        ::mdet::set_error(__VA_ARGS__); \
        return (code);                  \
    } while (0)
