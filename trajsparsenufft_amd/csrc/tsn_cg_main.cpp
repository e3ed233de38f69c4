// tsn_cg: command-line CG-SENSE reconstruction on libtsn.so, a drop-in for
// the reference binary nufft_fp16x2_cg
// (experiments/production/src/nufft_fp16x2_cg_bench.cu:279-518):
//
//   tsn_cg FWD_REAL FWD_IMAG ADJ_REAL ADJ_IMAG RUNTIME_DIR OPERATOR_SCALE
//          [warmup=5] [iters=20] [stress=0] [graph=0] [cg_iterations=10] [dump_dir]
//
// Same arguments, runtime files, dump files, one-line JSON on stdout and exit
// codes: 0 ok, 1 runtime/file error, 2 usage / incompatible operator, 3 CG
// non-positive curvature or non-finite reconstruction.
//
// RUNTIME_DIR holds measurement.c64.bin (complex64 [coils][samples], raw
// scale), sensitivity_maps.c64.bin (complex64 [coils][N][N]) and
// scaling.f32.bin (float32 [N][N], the apodization).  Shapes come from the
// packs: grid K x K = forward n_cols, image N = K / 2, coils from the maps
// file size.  TSN_PRECISION selects split16 (default), dense16 or hi16.
// dump_dir receives reconstruction.c64.bin ([N][N]), residuals.f32.bin
// (sqrt(rho) per iteration, float32 as the reference writes it) and
// residuals.f64.bin (the float64 history this library keeps).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/tsn.h"

namespace {

struct Failure {
  int code;
};

void check_tsn(int rc, const char* what) {
  if (rc != TSN_OK) {
    std::fprintf(stderr, "%s failed: %s\n", what, tsn_last_error());
    throw Failure{rc == TSN_ERR_ARGUMENT ? 2 : rc};
  }
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    throw Failure{1};
  }
}

template <typename T>
std::vector<T> read_file(const std::string& path, size_t count) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path.c_str());
    throw Failure{1};
  }
  const size_t bytes = static_cast<size_t>(f.tellg());
  if (count == 0) count = bytes / sizeof(T);
  if (bytes != count * sizeof(T)) {
    std::fprintf(stderr, "%s holds %zu bytes, expected %zu\n", path.c_str(), bytes, count * sizeof(T));
    throw Failure{1};
  }
  std::vector<T> v(count);
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(bytes));
  if (!f) {
    std::fprintf(stderr, "short read of %s\n", path.c_str());
    throw Failure{1};
  }
  return v;
}

template <typename T>
void write_file(const std::string& path, const T* data, size_t count) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(count * sizeof(T)));
  if (!f) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    throw Failure{1};
  }
}

template <typename T>
T* device_alloc(size_t count, std::vector<void*>& owned) {
  void* p = nullptr;
  check_hip(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
  owned.push_back(p);
  return static_cast<T*>(p);
}

int precision_from_env(const char** variant) {
  const char* env = std::getenv("TSN_PRECISION");
  const std::string p = env ? env : "split16";
  if (p == "split16") {
    *variant = "fp16x2_cg10";
    return TSN_PREC_SPLIT16;
  }
  if (p == "dense16") {
    *variant = "fp16x2_dense_tc_cg10";
    return TSN_PREC_DENSE16;
  }
  if (p == "hi16") {
    *variant = "fp16_hi_cg10";
    return TSN_PREC_HI16;
  }
  std::fprintf(stderr, "TSN_PRECISION must be split16, dense16 or hi16 (packs carry FP16 tiles)\n");
  throw Failure{2};
}

int run(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr,
                 "usage: %s FWD_REAL FWD_IMAG ADJ_REAL ADJ_IMAG RUNTIME_DIR OPERATOR_SCALE [warmup=5] "
                 "[iters=20] [stress=0] [graph=0] [cg_iterations=10] [dump_dir]\n",
                 argv[0]);
    return 2;
  }
  const std::string runtime = argv[5];
  const float operator_scale = std::strtof(argv[6], nullptr);
  const int warmup = argc > 7 ? std::atoi(argv[7]) : 5;
  const int iterations = argc > 8 ? std::atoi(argv[8]) : 20;
  const int stress = argc > 9 ? std::atoi(argv[9]) : 0;
  const bool graph_mode = argc > 10 ? std::atoi(argv[10]) != 0 : false;
  const int cg_iterations = argc > 11 ? std::atoi(argv[11]) : 10;
  const bool dump = argc > 12;
  const std::string dump_dir = dump ? argv[12] : "";
  if (!(operator_scale > 0.0f) || !std::isfinite(operator_scale) || cg_iterations <= 0 || iterations <= 0 ||
      warmup < 0 || stress < 0)
    return 2;
  const char* variant = nullptr;
  const int precision = precision_from_env(&variant);

  tsn_direction *fwd = nullptr, *adj = nullptr;
  check_tsn(tsn_direction_load(argv[1], argv[2], &fwd), "loading the forward pack");
  check_tsn(tsn_direction_load(argv[3], argv[4], &adj), "loading the adjoint pack");
  tsn_pack_header hf{}, ha{};
  check_tsn(tsn_direction_header(fwd, &hf), "forward header");
  check_tsn(tsn_direction_header(adj, &ha), "adjoint header");
  const uint32_t grid = static_cast<uint32_t>(std::lround(std::sqrt(static_cast<double>(hf.n_cols))));
  if (grid * grid != hf.n_cols || grid % 2 || ha.output_size != hf.n_cols || hf.output_size != ha.n_cols) {
    std::fprintf(stderr, "unexpected operator shape\n");
    return 2;
  }
  const uint32_t image = grid / 2;
  const size_t pixels = static_cast<size_t>(image) * image;
  const auto maps = read_file<float>(runtime + "/sensitivity_maps.c64.bin", 0);
  if (maps.size() % (2 * pixels) != 0 || maps.empty()) {
    std::fprintf(stderr, "sensitivity maps do not match a %ux%u image\n", image, image);
    return 2;
  }
  const uint32_t coils = static_cast<uint32_t>(maps.size() / (2 * pixels));
  const size_t samples = static_cast<size_t>(coils) * hf.output_size;
  const auto meas = read_file<float>(runtime + "/measurement.c64.bin", 2 * samples);
  const auto apod = read_file<float>(runtime + "/scaling.f32.bin", pixels);

  std::vector<void*> owned;
  float* d_maps = device_alloc<float>(maps.size(), owned);
  float* d_apod = device_alloc<float>(pixels, owned);
  float* d_meas = device_alloc<float>(2 * samples, owned);
  float* d_x = device_alloc<float>(2 * pixels, owned);
  double* d_hist = device_alloc<double>(cg_iterations + 1, owned);
  int* d_status = device_alloc<int>(1, owned);
  check_hip(hipMemcpy(d_maps, maps.data(), maps.size() * sizeof(float), hipMemcpyHostToDevice), "copy maps");
  check_hip(hipMemcpy(d_apod, apod.data(), pixels * sizeof(float), hipMemcpyHostToDevice), "copy scaling");
  check_hip(hipMemcpy(d_meas, meas.data(), 2 * samples * sizeof(float), hipMemcpyHostToDevice), "copy data");

  tsn_plan_config cfg{image, image, grid, grid, coils, precision};
  tsn_plan* plan = nullptr;
  check_tsn(tsn_plan_create(&cfg, fwd, adj, d_maps, d_apod, &plan), "plan");
  hipStream_t stream = nullptr;
  check_hip(hipStreamCreate(&stream), "stream");

  check_tsn(tsn_cg(plan, d_meas, static_cast<uint32_t>(cg_iterations), operator_scale, 1e-4f, d_x, d_hist,
                   d_status, stream),
            "cg");
  check_hip(hipStreamSynchronize(stream), "cg");
  int status = 0;
  std::vector<double> hist(cg_iterations + 1);
  std::vector<float> x(2 * pixels);
  check_hip(hipMemcpy(&status, d_status, sizeof(int), hipMemcpyDeviceToHost), "status");
  check_hip(hipMemcpy(hist.data(), d_hist, hist.size() * sizeof(double), hipMemcpyDeviceToHost), "history");
  check_hip(hipMemcpy(x.data(), d_x, x.size() * sizeof(float), hipMemcpyDeviceToHost), "image");
  size_t nonfinite = 0;
  for (size_t i = 0; i < pixels; ++i)
    if (!std::isfinite(x[2 * i]) || !std::isfinite(x[2 * i + 1])) ++nonfinite;
  if (status != 1 || nonfinite != 0) {
    std::fprintf(stderr, "CG correctness failure curvature=%d nonfinite=%zu\n", status, nonfinite);
    return 3;
  }

  if (graph_mode)
    check_tsn(tsn_cg_graph_capture(plan, d_meas, static_cast<uint32_t>(cg_iterations), operator_scale, 1e-4f, d_x,
                                   d_hist, d_status, stream),
              "graph capture");
  auto launch = [&]() {
    if (graph_mode)
      check_tsn(tsn_cg_graph_launch(plan, stream), "graph launch");
    else
      check_tsn(tsn_cg(plan, d_meas, static_cast<uint32_t>(cg_iterations), operator_scale, 1e-4f, d_x, d_hist,
                       d_status, stream),
                "cg");
  };
  for (int i = 0; i < warmup; ++i) launch();
  check_hip(hipStreamSynchronize(stream), "warmup");
  hipEvent_t start = nullptr, stop = nullptr;
  check_hip(hipEventCreate(&start), "event");
  check_hip(hipEventCreate(&stop), "event");
  check_hip(hipEventRecord(start, stream), "event");
  for (int i = 0; i < iterations; ++i) launch();
  check_hip(hipEventRecord(stop, stream), "event");
  check_hip(hipEventSynchronize(stop), "event");
  float elapsed_ms = 0.0f;
  check_hip(hipEventElapsedTime(&elapsed_ms, start, stop), "event");
  for (int i = 0; i < stress; ++i) launch();
  check_hip(hipStreamSynchronize(stream), "stress");

  if (dump) {
    check_hip(hipMemcpy(x.data(), d_x, x.size() * sizeof(float), hipMemcpyDeviceToHost), "image");
    check_hip(hipMemcpy(hist.data(), d_hist, hist.size() * sizeof(double), hipMemcpyDeviceToHost), "history");
    std::vector<float> hist32(hist.begin(), hist.end());
    write_file(dump_dir + "/reconstruction.c64.bin", x.data(), x.size());
    write_file(dump_dir + "/residuals.f32.bin", hist32.data(), hist32.size());
    write_file(dump_dir + "/residuals.f64.bin", hist.data(), hist.size());
  }
  std::printf(
      "{\"variant\":\"%s\",\"correct\":true,\"us\":%.9g,\"cg_iterations\":%d,\"operator_scale\":%.9g,"
      "\"residual0\":%.9g,\"residual_final\":%.9g,\"graph\":%s,\"telemetry\":false,\"nonfinite\":0}\n",
      variant, static_cast<double>(elapsed_ms) * 1000.0 / iterations, cg_iterations,
      static_cast<double>(operator_scale), hist.front(), hist.back(), graph_mode ? "true" : "false");

  (void)hipEventDestroy(start);
  (void)hipEventDestroy(stop);
  (void)hipStreamDestroy(stream);
  tsn_plan_free(plan);
  for (void* p : owned) (void)hipFree(p);
  tsn_direction_free(fwd);
  tsn_direction_free(adj);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const Failure& f) {
    return f.code;
  }
}
