// Readers and writer of the drpt driver (see io.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace drpt_io {

int read_idx3_ubyte(const std::string& path, int64_t n, int d, std::vector<uint8_t>& out,
                    std::string& err);
int read_fbin(const std::string& path, int64_t n, int d, int bytes_per_value, int64_t offset,
              std::vector<float>& out, std::string& err);
// idx: n x K neighbour lists (-1 padded); writes the first nn_write non-self entries per source
int write_edge_list(const std::string& path, const int32_t* idx, int64_t n, int K, int nn_write,
                    std::string& err);

}  // namespace drpt_io
