// Input readers and the edge-list writer of the drpt driver.
//
// Formats kept from the reference:
//  * idx3-ubyte (ImageReader::read_ubyte, cpp/src/drpt/io/file_reader.cpp:33-87):
//    four big-endian int32 (magic, count, rows, cols), then count*rows*cols
//    bytes, one image per row. The magic number is not checked (as in the
//    reference). Rows are taken as stored (the reference's arr[i][rows*r + c]
//    indexing equals this for square images and for rows == 1).
//  * fbin (ImageReader::mpi_file_read, file_reader.cpp:364-493): a header of
//    `offset` bytes, then float32 (bytes_per_value == 4) rows of `dimension`
//    values. Unlike the reference (which replaces every value by a random
//    integer, file_reader.cpp:476-485), the stored values are used.
//  * edge list (FileWriter<int>::mpi_write_edge_list, file_writer.cpp:36-184):
//    for each source in ascending order, its first min(nn, len) neighbours
//    other than itself, one "src+1 dst+1\n" line each. The file is truncated
//    before writing (the reference opens without truncation).
#include "io.hpp"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>

namespace drpt_io {

namespace {
uint32_t be32(const unsigned char* b) {
  return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
}
}  // namespace

int read_idx3_ubyte(const std::string& path, int64_t n, int d, std::vector<uint8_t>& out,
                    std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path + ": " + std::strerror(errno);
    return -1;
  }
  unsigned char hdr[16];
  if (!f.read((char*)hdr, 16)) {
    err = path + ": truncated idx3 header";
    return -1;
  }
  const int64_t count = (int32_t)be32(hdr + 4);
  const int64_t rows = (int32_t)be32(hdr + 8), cols = (int32_t)be32(hdr + 12);
  if (rows * cols != d) {
    err = path + ": image size " + std::to_string(rows) + "x" + std::to_string(cols) +
          " does not match -dimension " + std::to_string(d);
    return -1;
  }
  if (count < n) {
    err = path + ": holds " + std::to_string(count) + " images, fewer than -data-set-size " +
          std::to_string(n);
    return -1;
  }
  out.resize((size_t)n * d);
  if (!f.read((char*)out.data(), (std::streamsize)out.size())) {
    err = path + ": truncated idx3 payload";
    return -1;
  }
  return 0;
}

int read_fbin(const std::string& path, int64_t n, int d, int bytes_per_value, int64_t offset,
              std::vector<float>& out, std::string& err) {
  if (bytes_per_value != 4) {
    err = "fbin: only 4-byte float values are supported (-bytes-for-data-type 4)";
    return -1;
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path + ": " + std::strerror(errno);
    return -1;
  }
  f.seekg(0, std::ios::end);
  const int64_t size = (int64_t)f.tellg();
  const int64_t need = offset + n * (int64_t)d * 4;
  if (size < need) {
    err = path + ": " + std::to_string(size) + " bytes, need " + std::to_string(need);
    return -1;
  }
  f.seekg(offset, std::ios::beg);
  out.resize((size_t)n * d);
  if (!f.read((char*)out.data(), (std::streamsize)(out.size() * 4))) {
    err = path + ": read failed";
    return -1;
  }
  return 0;
}

int write_edge_list(const std::string& path, const int32_t* idx, int64_t n, int K, int nn_write,
                    std::string& err) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path + " for writing: " + std::strerror(errno);
    return -1;
  }
  std::vector<char> buf;
  buf.reserve(1 << 20);
  char line[32];
  for (int64_t s = 0; s < n; ++s) {
    int written = 0;
    for (int k = 0; k < K && written < nn_write; ++k) {
      const int32_t dst = idx[s * K + k];
      if (dst < 0) break;
      if (dst == s) continue;
      const int len = std::snprintf(line, sizeof line, "%lld %lld\n", (long long)s + 1,
                                    (long long)dst + 1);
      buf.insert(buf.end(), line, line + len);
      ++written;
    }
    if (buf.size() > (1u << 20)) {
      std::fwrite(buf.data(), 1, buf.size(), f);
      buf.clear();
    }
  }
  std::fwrite(buf.data(), 1, buf.size(), f);
  if (std::fclose(f) != 0) {
    err = "write to " + path + " failed";
    return -1;
  }
  return 0;
}

}  // namespace drpt_io
