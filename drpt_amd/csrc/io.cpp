// Input readers, host-to-device upload and the edge-list writer of the drpt
// driver.
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
//
// Speed: the reference reads one byte per ifstream::read (file_reader.cpp:65-83)
// and formats with one sprintf per number (file_writer.cpp:141-150). Here the
// inputs are memory-mapped and copied to the device through two pinned
// staging buffers (host copies by several threads overlapping the
// asynchronous DMA of the previous buffer), and the edge list is formatted by
// several threads into per-range buffers written at their prefix-summed file
// offsets (pwrite).
#include "io.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <thread>
#include <vector>

namespace drpt_io {

namespace {

uint32_t be32(const unsigned char* b) {
  return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
}

int map_file(const std::string& path, Mapped& m, std::string& err) {
  const int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) {
    err = "cannot open " + path + ": " + std::strerror(errno);
    return -1;
  }
  struct stat st;
  if (::fstat(fd, &st) != 0) {
    err = "cannot stat " + path + ": " + std::strerror(errno);
    ::close(fd);
    return -1;
  }
  m.size = (size_t)st.st_size;
  if (m.size > 0) {
    void* p = ::mmap(nullptr, m.size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
      err = "cannot map " + path + ": " + std::strerror(errno);
      ::close(fd);
      m.size = 0;
      return -1;
    }
    (void)::madvise(p, m.size, MADV_SEQUENTIAL);
    (void)::madvise(p, m.size, MADV_WILLNEED);
    m.data = static_cast<const unsigned char*>(p);
  }
  ::close(fd);  // the mapping stays valid
  return 0;
}

// runs f(i) for i in [0, parts) on up to `threads` threads
template <typename F>
void parallel_for(int parts, int threads, F&& f) {
  threads = std::max(1, std::min(threads, parts));
  if (threads == 1) {
    for (int i = 0; i < parts; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      for (int i = t; i < parts; i += threads) f(i);
    });
  for (auto& th : pool) th.join();
}

// decimal digits of v > 0 written at p, returns the end
char* put_uint(char* p, uint64_t v) {
  char tmp[24];
  int k = 0;
  do {
    tmp[k++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (k) *p++ = tmp[--k];
  return p;
}

}  // namespace

Mapped::~Mapped() {
  if (data) ::munmap(const_cast<unsigned char*>(data), size);
}

int open_idx3_ubyte(const std::string& path, int64_t n, int d, Mapped& m,
                    const uint8_t** payload, std::string& err) {
  if (map_file(path, m, err)) return -1;
  if (m.size < 16) {
    err = path + ": truncated idx3 header";
    return -1;
  }
  const unsigned char* hdr = m.data;
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
  if ((int64_t)m.size < 16 + n * (int64_t)d) {
    err = path + ": truncated idx3 payload";
    return -1;
  }
  *payload = m.data + 16;
  return 0;
}

int open_fbin(const std::string& path, int64_t n, int d, int bytes_per_value, int64_t offset,
              Mapped& m, const float** payload, std::string& err) {
  if (bytes_per_value != 4) {
    err = "fbin: only 4-byte float values are supported (-bytes-for-data-type 4)";
    return -1;
  }
  if (map_file(path, m, err)) return -1;
  const int64_t need = offset + n * (int64_t)d * 4;
  if ((int64_t)m.size < need) {
    err = path + ": " + std::to_string(m.size) + " bytes, need " + std::to_string(need);
    return -1;
  }
  if (offset % 4 != 0) {
    err = path + ": the header length (-data-starting-index) must be a multiple of 4";
    return -1;
  }
  *payload = reinterpret_cast<const float*>(m.data + offset);
  return 0;
}

int upload(const void* src, size_t bytes, void* dst_dev, hipStream_t stream, int threads,
           std::string& err) {
  constexpr size_t kChunk = (size_t)64 << 20;
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  auto fail = [&](const char* what, hipError_t e) {
    err = std::string(what) + " failed: " + hipGetErrorString(e);
    return -1;
  };
  int rc = 0;
  for (int b = 0; b < 2 && rc == 0; ++b) {
    hipError_t e = hipHostMalloc(&stage[b], std::min(kChunk, std::max<size_t>(bytes, 1)), 0);
    if (e != hipSuccess) rc = fail("hipHostMalloc", e);
    else if ((e = hipEventCreate(&done[b])) != hipSuccess) rc = fail("hipEventCreate", e);
  }
  const unsigned char* s = static_cast<const unsigned char*>(src);
  unsigned char* dst = static_cast<unsigned char*>(dst_dev);
  for (size_t off = 0, i = 0; rc == 0 && off < bytes; off += kChunk, ++i) {
    const size_t len = std::min(kChunk, bytes - off);
    const int b = (int)(i & 1);
    hipError_t e = hipEventSynchronize(done[b]);  // the buffer's previous copy has ended
    if (e != hipSuccess) {
      rc = fail("hipEventSynchronize", e);
      break;
    }
    unsigned char* st = static_cast<unsigned char*>(stage[b]);
    const int parts = std::max(1, threads * 2);
    const size_t piece = (len + parts - 1) / parts;
    parallel_for(parts, threads, [&](int p) {
      const size_t a = std::min(len, (size_t)p * piece), z = std::min(len, a + piece);
      if (z > a) std::memcpy(st + a, s + off + a, z - a);
    });
    if ((e = hipMemcpyAsync(dst + off, st, len, hipMemcpyHostToDevice, stream)) != hipSuccess)
      rc = fail("hipMemcpyAsync", e);
    else if ((e = hipEventRecord(done[b], stream)) != hipSuccess)
      rc = fail("hipEventRecord", e);
  }
  if (rc == 0) {
    const hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) rc = fail("hipStreamSynchronize", e);
  } else {
    (void)hipStreamSynchronize(stream);
  }
  for (int b = 0; b < 2; ++b) {
    if (done[b]) (void)hipEventDestroy(done[b]);
    if (stage[b]) (void)hipHostFree(stage[b]);
  }
  return rc;
}

int write_edge_list(const std::string& path, const int32_t* idx, int64_t n, int K, int nn_write,
                    int threads, std::string& err) {
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) {
    err = "cannot open " + path + " for writing: " + std::strerror(errno);
    return -1;
  }
  // ranges of sources formatted independently (at most 2^20 sources each)
  const int64_t per = std::max<int64_t>(4096, std::min<int64_t>((int64_t)1 << 20,
                                                                (n + threads * 8 - 1) / (threads * 8)));
  const int parts = (int)((n + per - 1) / per);
  std::vector<std::vector<char>> buf(parts);
  parallel_for(parts, threads, [&](int p) {
    const int64_t s0 = (int64_t)p * per, s1 = std::min(n, s0 + per);
    std::vector<char>& b = buf[p];
    b.resize((size_t)(s1 - s0) * (size_t)std::max(nn_write, 0) * 24);
    char* w = b.data();
    for (int64_t s = s0; s < s1; ++s) {
      int written = 0;
      for (int k = 0; k < K && written < nn_write; ++k) {
        const int32_t dst = idx[s * K + k];
        if (dst < 0) break;
        if (dst == s) continue;
        w = put_uint(w, (uint64_t)s + 1);
        *w++ = ' ';
        w = put_uint(w, (uint64_t)dst + 1);
        *w++ = '\n';
        ++written;
      }
    }
    b.resize((size_t)(w - b.data()));
  });
  std::vector<off_t> at(parts + 1, 0);
  for (int p = 0; p < parts; ++p) at[p + 1] = at[p] + (off_t)buf[p].size();
  std::vector<int> bad(parts, 0);
  parallel_for(parts, threads, [&](int p) {
    const char* q = buf[p].data();
    size_t left = buf[p].size();
    off_t o = at[p];
    while (left > 0) {
      const ssize_t r = ::pwrite(fd, q, left, o);
      if (r <= 0) {
        bad[p] = errno ? errno : EIO;
        return;
      }
      q += r;
      left -= (size_t)r;
      o += r;
    }
  });
  int rc = 0;
  for (int p = 0; p < parts; ++p)
    if (bad[p]) {
      err = "write to " + path + " failed: " + std::strerror(bad[p]);
      rc = -1;
      break;
    }
  if (::close(fd) != 0 && rc == 0) {
    err = "write to " + path + " failed: " + std::strerror(errno);
    rc = -1;
  }
  return rc;
}

}  // namespace drpt_io
