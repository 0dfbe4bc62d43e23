// Compile-time tables of the TNF row layout, shared by host and device code.
#pragma once
#include <cstdint>

namespace mdet {

// Column order of the 136 canonical tetranucleotides: the TNF row layout of
// the reference (openmp/TNF.cpp:14-24, TN[]).  Each is the lexicographically
// smaller of a 4-mer and its reverse complement.
constexpr char kTnColumnsCx[136 * 4 + 1] =
    "GGTAAGCCAAAAACATAGTCACGACATACGAAAAGTCAAACCAGGGACATTAGATCCCTCCTAAACTAAGGC"
    "GCAACCGCCGCCAAACACTCATCCGACCGAGAATAGATCACAGAAGTAATGAAAATTTAATATAAGTGAGCT"
    "CCACGGCCACCCGGGAGCGCATACCTGATAGAATATGTCACTCCACAAACCTTAAAAACGCGAGAGGGATCG"
    "ACGCTCAACTACCTCAGACAGGAACTTCGCCCCTGCTGCAGGCACACGGAGCAACTCATGAATTACAGAGAT"
    "ATAACATCGCCATCGACACACAACAAGGAGCAATGGATTCGTGAACCGGATAGCTACGTCCCCGAAGCCGTA"
    "GTACAGGAAATGCACCCAGCCGGCACACCCGGCCGACCCCTGAAAACAAGAGCCCACGGATACAACCAACGT"
    "GAACGTAAATGCGTTATCCACAGGACTGAAAGAAGACAAGGCGAAACCACGGCCAACTTAAGACAGCGGAAA"
    "AATCATTGGCACCCTACGACCTAGAGAACGCACGCGAATA";

constexpr int base_code_cx(char c) { return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : -1; }

// 2-bit base code: A=0 C=1 G=2 T=3 (complement = 3 - b).  A 4-mer code puts
// the first base in the low bits: code = b0 | b1<<2 | b2<<4 | b3<<6.
struct ColumnMap {
    int16_t col[256];
};

constexpr int code_of_string_cx(const char* s) {
    return base_code_cx(s[0]) | (base_code_cx(s[1]) << 2) | (base_code_cx(s[2]) << 4) | (base_code_cx(s[3]) << 6);
}
constexpr int revcomp_code_cx(int code) {
    int r = 0;
    for (int k = 0; k < 4; ++k) r |= (3 - ((code >> (2 * k)) & 3)) << (2 * (3 - k));
    return r;
}
constexpr ColumnMap make_column_map() {
    ColumnMap m{};
    for (int i = 0; i < 256; ++i) m.col[i] = -1;
    for (int c = 0; c < 136; ++c) {
        int code = code_of_string_cx(kTnColumnsCx + 4 * c);
        m.col[code] = (int16_t)c;
        m.col[revcomp_code_cx(code)] = (int16_t)c;
    }
    return m;
}
constexpr ColumnMap kColumnMap = make_column_map();

constexpr bool column_map_complete() {
    for (int i = 0; i < 256; ++i)
        if (kColumnMap.col[i] < 0) return false;
    return true;
}
static_assert(column_map_complete(), "every 4-mer must map to a TNF column");

}  // namespace mdet
