// Host STL parsing shared by the host loader (s2hip_host.cpp) and the device
// mesh ingest (s2hip_meshin.hip).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace s2h {
namespace stlio {
bool read_file(const char *path, std::vector<unsigned char> &out);
// looks_ascii_stl (src/io/stl.rs:201-212)
bool looks_ascii(const std::vector<unsigned char> &b);
// ASCII "vertex x y z" corners, three per face (src/io/stl.rs:82-153); false
// when no face was found (the caller then parses the bytes as binary).
bool ascii_corners(const std::vector<unsigned char> &b, std::vector<double> &xyz);
// Binary face count, checked against the length (src/io/stl.rs:157-189).
bool binary_count(const unsigned char *b, size_t n, uint32_t &count, std::string &err);
}  // namespace stlio
}  // namespace s2h
