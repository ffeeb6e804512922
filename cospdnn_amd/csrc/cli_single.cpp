// cospdnn.out <neuron> <batch> <layer> <gpu_num> <gpu_id> <dataset_dir> [options]
// Drop-in for the reference's single-task driver (src/cospdnn.cpp:80-188).
// Options: --precision=fp16|fp32  --guard-layer=D  --graph  --no-fuse  --bias=B  --categories=FILE
#include <hip/hip_runtime.h>

#include <fstream>
#include <thread>

#include "cli_common.hpp"

int main(int argc, char* argv[]) {
  if (argc < 7) {
    std::cout << "Usage: exe neuron batch layer gpu_num gpu_id dataset_dir" << std::endl;
    return 0;
  }
  const int neuron = std::atoi(argv[1]);
  const int batch = std::atoi(argv[2]);
  const int layer = std::atoi(argv[3]);
  int gpu_num = std::atoi(argv[4]);
  const int dev = std::atoi(argv[5]);
  const std::string dir(argv[6]);
  cli::Flags f;
  if (!cli::parse_flags(argc, argv, 7, f)) return 2;
  const float bias = f.bias_set ? f.bias : cli::bias_for(neuron);

  std::cout << "[BEGIN]..." << std::endl;
  std::cout << "[Batch] :" << batch << std::endl;
  std::cout << "[Neuron] :" << neuron << std::endl;
  std::cout << "[Layer] :" << layer << std::endl;
  cli::Dataset d;
  cli::load_dataset(dir, neuron, batch, layer, f.use_fuse, d);

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "no HIP device\n");
    return EXIT_FAILURE;
  }
  if (gpu_num < 1) gpu_num = 1;
  if (gpu_num > ndev) {
    std::cout << "Exceeding Maximum GPU Number! Automatically Change to Max GPU Number " << ndev << std::endl;
    gpu_num = ndev;
  }
  // one host thread per GPU, each on its slice of ceil(batch/gpu_num) samples (cospdnn.cpp:155-186)
  const int per = (batch + gpu_num - 1) / gpu_num;
  std::vector<std::vector<int64_t>> cats(gpu_num);
  std::vector<double> ms(gpu_num, 0.0);
  std::vector<int> rcs(gpu_num, COSPDNN_OK);
  auto work = [&](int g) {
    const int device = gpu_num == 1 ? dev : g;
    const int64_t s0 = static_cast<int64_t>(g) * per;
    const int64_t s1 = std::min<int64_t>(s0 + per, batch);
    if (s1 <= s0) return;
    cospdnn_engine* eng = nullptr;
    cospdnn_net* net = nullptr;
    int rc = cospdnn_engine_create(device, &eng);
    if (rc == COSPDNN_OK) rc = cli::make_net(eng, d, s0, s1, bias, f, &net);
    if (rc == COSPDNN_OK) {
      cats[g].assign(s1 - s0, 0);
      cospdnn_result r;
      std::memset(&r, 0, sizeof(r));
      r.categories = cats[g].data();
      const cospdnn_opts o = cli::run_opts(f);
      rc = cospdnn_run(&net, 1, &o, &r);
      if (rc == COSPDNN_OK) {
        cats[g].resize(r.n_categories);
        for (auto& c : cats[g]) c += s0;
        ms[g] = r.ms_total;
      }
    }
    cospdnn_net_destroy(net);
    cospdnn_engine_destroy(eng);
    rcs[g] = rc;
  };
  if (gpu_num == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int g = 0; g < gpu_num; ++g) th.emplace_back(work, g);
    for (auto& t : th) t.join();
  }
  int64_t total = 0;
  for (int g = 0; g < gpu_num; ++g) {
    if (rcs[g] != COSPDNN_OK) cli::die(rcs[g]);
    if (gpu_num > 1) std::cout << "GPU[" << g << "] ";
    std::cout << "Kernel Exec Time = " << ms[g] << "ms" << std::endl;
    total += static_cast<int64_t>(cats[g].size());
  }
  std::cout << "categories = " << total << std::endl;
  if (!f.categories.empty()) {
    std::ofstream out(f.categories);
    for (auto& c : cats)
      for (int64_t id : c) out << (id + 1) << "\n";
  }
  std::cout << "[END]..." << std::endl;
  return 0;
}
