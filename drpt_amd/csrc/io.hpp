// Readers, upload and edge-list writer of the drpt driver (see io.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace drpt_io {

// A read-only memory mapping of a whole file.
struct Mapped {
  const unsigned char* data = nullptr;
  size_t size = 0;
  Mapped() = default;
  Mapped(const Mapped&) = delete;
  Mapped& operator=(const Mapped&) = delete;
  ~Mapped();
};

// idx3-ubyte: maps the file, checks the header, points `payload` at the first
// n images (n * d bytes).
int open_idx3_ubyte(const std::string& path, int64_t n, int d, Mapped& m,
                    const uint8_t** payload, std::string& err);
// fbin: maps the file, points `payload` at n rows of d float32 after `offset`
// header bytes.
int open_fbin(const std::string& path, int64_t n, int d, int bytes_per_value, int64_t offset,
              Mapped& m, const float** payload, std::string& err);
// Copies bytes from host memory (a mapping) to device memory through two
// pinned staging buffers: `threads` threads fill one buffer while the other's
// asynchronous copy runs on `stream`. Returns 0 or -1 (err set).
int upload(const void* src, size_t bytes, void* dst_dev, hipStream_t stream, int threads,
           std::string& err);
// idx: n x K neighbour lists (-1 padded); writes, for each source in ascending
// order, its first nn_write entries other than itself as "src+1 dst+1\n",
// formatted by `threads` threads and written at their offsets.
int write_edge_list(const std::string& path, const int32_t* idx, int64_t n, int K, int nn_write,
                    int threads, std::string& err);

}  // namespace drpt_io
