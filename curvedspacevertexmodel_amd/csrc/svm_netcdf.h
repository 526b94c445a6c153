// Host-side writer of the reference's trajectory records
// (vectorValueNetCDF, src/databases/vectorValueNetCDF.cpp:3-36 with
// baseNetCDFDatabase, src/databases/baseNetCDFDatabase.cpp:3-79): a NetCDF
// classic-format file (CDF-1, big-endian) with dimensions record (unlimited),
// unit = 1 and N, and the double variables value(record, unit) and
// vector(record, N); each writeState appends one record and updates the
// record count in the header.  The NetCDF library is not needed: the classic
// format is a fixed published layout, written here directly.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct svm_nc_file {
    FILE* f = nullptr;
    int32_t n = 0;
    int32_t numrecs = 0;
    long header = 0;     // bytes before the first record
    long recsize = 0;    // 8 (value) + 8 n (vector)
};

namespace svm_nc {

inline void be32(std::string& b, uint32_t x) {
    for (int k = 3; k >= 0; --k) b.push_back((char)((x >> (8 * k)) & 0xff));
}
inline void name(std::string& b, const char* s) {
    const uint32_t len = (uint32_t)std::strlen(s);
    be32(b, len);
    b.append(s, len);
    while (b.size() % 4) b.push_back('\0');
}
inline void be64d(unsigned char* out, double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    for (int k = 7; k >= 0; --k) out[7 - k] = (unsigned char)((u >> (8 * k)) & 0xff);
}

constexpr uint32_t NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_DOUBLE = 6;

// The header for begin offsets b_value, b_vector (their size does not depend on them).
inline std::string header(int32_t n, uint32_t b_value, uint32_t b_vector) {
    std::string h("CDF\x01", 4);
    be32(h, 0);                                    // numrecs, rewritten by every record
    be32(h, NC_DIMENSION);
    be32(h, 3);
    name(h, "record"); be32(h, 0);                 // the unlimited dimension
    name(h, "unit");   be32(h, 1);
    name(h, "N");      be32(h, (uint32_t)n);
    be32(h, 0); be32(h, 0);                        // no global attributes
    be32(h, NC_VARIABLE);
    be32(h, 2);
    name(h, "value");
    be32(h, 2); be32(h, 0); be32(h, 1);            // (record, unit)
    be32(h, 0); be32(h, 0);                        // no attributes
    be32(h, NC_DOUBLE); be32(h, 8); be32(h, b_value);
    name(h, "vector");
    be32(h, 2); be32(h, 0); be32(h, 2);            // (record, N)
    be32(h, 0); be32(h, 0);
    be32(h, NC_DOUBLE); be32(h, (uint32_t)(8 * (int64_t)n)); be32(h, b_vector);
    return h;
}

}  // namespace svm_nc
