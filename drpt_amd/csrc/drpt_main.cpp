// drpt — command-line driver of the MI355X k-NN graph builder.
//
// Keeps the reference driver's options and flow (cpp/src/main.cpp:25-259):
// parse flags (:43-97), validate (:99-127), default density 1/sqrt(d)
// (:129-132), tree depth int(log2(N) - 2) (:141-145), read (:158-180), build
// (:198-217), write nn-1 edges per point to <output>results.txt (:227-228) and
// print "IO Time (s) ... Index building (s) ... Querying Time (s) ..."
// (:247-251). One process drives one GPU; the build runs through the C ABI.
//
// Options beyond the reference: -seed (the reference draws its seeds from
// std::random_device), -precision fp32|bf16|fp8, -kprime, -leaf-max (split
// nodes larger than this instead of the reference local stage), -depth,
// -device, -tree-group (trees per pooled filter pass of a bf16/fp8 build),
// -autotune 1 (tune the work-item plan on this input before the build).
// -locality 0|1 (main.cpp:39; default 0 here): bf16/fp8 builds' row merges
// read a copy of the input in tree 0's leaf order (measured slower on one GPU:
// the rows' L2 reuse does not depend on their addresses, and the extra index
// load lengthens every gather).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/drpt_hip.h"
#include "io.hpp"

namespace {

void die(const char* msg) {
  std::printf("%s\n", msg);
  std::exit(1);
}

int check(int rc, const char* what) {
  if (rc != DRPT_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, drpt_last_error());
    std::exit(1);
  }
  return rc;
}

}  // namespace

int main(int argc, char* argv[]) {
  std::string input_path, output_path;
  long data_set_size = 0;
  int dimension = 0, ntrees = 10, nn = 0, local_tree_offset = 2, file_format = 0;
  int data_starting_index = 8, bytes_for_data_type = 4;
  double tree_depth_ratio = 0.6, density = 0;
  uint32_t seed = 1;
  std::string precision = "fp32";
  int kprime = 0, leaf_max = 0, depth_override = 0, device = 0, tree_group = 0, autotune = 0;
  int locality = 0;
  for (int p = 0; p + 1 < argc; p++) {
    const char* a = argv[p];
    const char* v = argv[p + 1];
    if (!std::strcmp(a, "-input")) input_path = v;
    else if (!std::strcmp(a, "-output")) output_path = v;
    else if (!std::strcmp(a, "-data-set-size")) data_set_size = std::atol(v);
    else if (!std::strcmp(a, "-dimension")) dimension = std::atoi(v);
    else if (!std::strcmp(a, "-ntrees")) ntrees = std::atoi(v);
    else if (!std::strcmp(a, "-tree-depth-ratio")) tree_depth_ratio = std::stof(v);
    else if (!std::strcmp(a, "-density")) density = std::atoi(v);  // integer, as main.cpp:71
    else if (!std::strcmp(a, "-nn")) nn = std::atoi(v);
    else if (!std::strcmp(a, "-locality")) locality = std::atoi(v);
    else if (!std::strcmp(a, "-local-tree-offset")) local_tree_offset = std::atoi(v);
    else if (!std::strcmp(a, "-data-file-format")) file_format = std::atoi(v);
    else if (!std::strcmp(a, "-bytes-for-data-type")) bytes_for_data_type = std::atoi(v);
    else if (!std::strcmp(a, "-data-starting-index")) data_starting_index = std::atoi(v);
    else if (!std::strcmp(a, "-seed")) seed = (uint32_t)std::strtoul(v, nullptr, 10);
    else if (!std::strcmp(a, "-precision")) precision = v;
    else if (!std::strcmp(a, "-kprime")) kprime = std::atoi(v);
    else if (!std::strcmp(a, "-leaf-max")) leaf_max = std::atoi(v);
    else if (!std::strcmp(a, "-depth")) depth_override = std::atoi(v);
    else if (!std::strcmp(a, "-device")) device = std::atoi(v);
    else if (!std::strcmp(a, "-tree-group")) tree_group = std::atoi(v);
    else if (!std::strcmp(a, "-autotune")) autotune = std::atoi(v);
  }
  if (input_path.empty()) die("Valid input path needed!...");
  if (output_path.empty()) die("Valid out path needed!...");
  if (data_set_size == 0) die("Dataset size should be greater than 0");
  if (dimension == 0) die("Dimension size should be greater than 0");
  if (nn == 0) die("Nearest neighbours size should be greater than 0");
  if (density == 0) density = 1.0 / std::sqrt((double)dimension);
  if (file_format != 0 && file_format != 1)
    die("Unsupported data file format, ubyte and bin file formats are supported!...");

  drpt_params p;
  check(drpt_derive_params(data_set_size, dimension, ntrees, nn, tree_depth_ratio, &p),
        "derive_params");
  std::printf(" tree depth %d\n", (int)(std::log2((double)data_set_size) - 2));
  p.local_offset = local_tree_offset;
  p.density = (float)density;
  p.seed = seed;
  p.kprime = kprime;
  p.tree_group = tree_group;
  p.locality = locality != 0 ? 1 : 0;
  if (precision == "fp32") p.precision = DRPT_FP32;
  else if (precision == "bf16") p.precision = DRPT_BF16;
  else if (precision == "fp8") p.precision = DRPT_FP8;
  else die("-precision must be fp32, bf16 or fp8");
  if (leaf_max > 0) {
    p.local_mode = DRPT_LOCAL_LEAFMAX;
    p.leaf_max = leaf_max;
    int need = (int)std::ceil(std::log2(std::max(2.0, (double)data_set_size / leaf_max))) + 4;
    if (need > p.depth) p.depth = need;
  }
  if (depth_override > 0) p.depth = depth_override;

  const int io_threads =
      (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // read: the file mapped, copied to the device through pinned staging buffers
  auto t_io0 = std::chrono::steady_clock::now();
  drpt_io::Mapped in;
  const void* payload = nullptr;
  std::string err;
  if (file_format == 0) {
    const uint8_t* pu = nullptr;
    if (drpt_io::open_idx3_ubyte(input_path, data_set_size, dimension, in, &pu, err))
      die(err.c_str());
    payload = pu;
    p.dtype = DRPT_U8;
  } else {
    const float* pf = nullptr;
    if (drpt_io::open_fbin(input_path, data_set_size, dimension, bytes_for_data_type,
                           data_starting_index, in, &pf, err))
      die(err.c_str());
    payload = pf;
    p.dtype = DRPT_F32;
  }
  drpt_ctx* ctx = nullptr;
  check(drpt_create(&ctx, device), "drpt_create");
  const size_t xbytes = (size_t)data_set_size * dimension * (p.dtype == DRPT_U8 ? 1 : 4);
  void* dX = nullptr;
  int32_t* dI = nullptr;
  float* dD = nullptr;
  if (hipMalloc(&dX, xbytes) != hipSuccess || hipMalloc((void**)&dI, 4ull * data_set_size * nn) ||
      hipMalloc((void**)&dD, 4ull * data_set_size * nn))
    die("device allocation failed");
  if (drpt_io::upload(payload, xbytes, dX, (hipStream_t)drpt_get_stream(ctx), io_threads, err))
    die(err.c_str());
  auto t_io1 = std::chrono::steady_clock::now();
  std::printf("Rank 0 IO reading completed %ld*%d\n", data_set_size, dimension);
  drpt_timings tm;
  if (autotune && p.precision != DRPT_FP32) {
    int32_t plan[11];
    check(drpt_autotune(ctx, &p, dX, plan), "drpt_autotune");
    std::printf("Work-item plan: rows per item %d %d %d %d, ring %d %d %d %d, trees per group %d, "
                "first trees %d, re-rank %d\n",
                plan[0], plan[1], plan[2], plan[3], plan[4], plan[5], plan[6], plan[7], plan[8],
                plan[9], plan[10]);
  }
  check(drpt_build_graph(ctx, &p, dX, dI, dD, &tm), "drpt_build_graph");
  auto t_w0 = std::chrono::steady_clock::now();
  int32_t* idx = nullptr;  // pinned: the lists come back at full PCIe rate
  if (hipHostMalloc((void**)&idx, 4ull * data_set_size * nn, 0) != hipSuccess ||
      hipMemcpy(idx, dI, 4ull * data_set_size * nn, hipMemcpyDeviceToHost) != hipSuccess)
    die("download failed");
  if (drpt_io::write_edge_list(output_path + "results.txt", idx, data_set_size, nn, nn - 1,
                               io_threads, err))
    die(err.c_str());
  (void)hipHostFree(idx);
  auto t_w1 = std::chrono::steady_clock::now();
  const double io_s = std::chrono::duration<double>(t_io1 - t_io0).count() +
                      std::chrono::duration<double>(t_w1 - t_w0).count();
  const double build_s = (tm.project_ms + tm.split_ms + tm.group_ms) / 1e3;
  const double query_s = (tm.knn_ms + tm.merge_ms) / 1e3;
  std::printf("IO Time (s)%g Index building (s) %g Querying Time (s) %g\n", io_s, build_s, query_s);
  std::printf("IO detail (s): read+upload %g, download+write %g (%d threads)\n",
              std::chrono::duration<double>(t_io1 - t_io0).count(),
              std::chrono::duration<double>(t_w1 - t_w0).count(), io_threads);
  std::ofstream stats(output_path + "stats.txt", std::ios::app);
  stats << "{\"n\": " << data_set_size << ", \"d\": " << dimension << ", \"ntrees\": " << ntrees
        << ", \"nn\": " << nn << ", \"precision\": \"" << precision << "\", \"project_ms\": "
        << tm.project_ms << ", \"split_ms\": " << tm.split_ms << ", \"group_ms\": " << tm.group_ms
        << ", \"knn_ms\": " << tm.knn_ms << ", \"merge_ms\": " << tm.merge_ms
        << ", \"total_ms\": " << tm.total_ms << ", \"leaves\": " << tm.n_leaves
        << ", \"fallback_rows\": " << tm.n_fallback_rows << "}\n";
  (void)hipFree(dX);
  (void)hipFree(dI);
  (void)hipFree(dD);
  drpt_destroy(ctx);
  return 0;
}
