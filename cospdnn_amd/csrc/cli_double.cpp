// cospdnn_double.out <neuron1> <batch1> <layer1> <neuron2> <batch2> <layer2> <gpu_id> <dataset_dir> [options]
// Drop-in for the reference's concurrent-task driver (src/cospdnn_double.cpp:81-197): two
// networks in one run on one GPU.  The reference requires equal layer counts
// (cospdnn_double.cu:243-248); --allow-uneven lifts that.
// Options: as cospdnn.out, plus --allow-uneven.
#include <algorithm>
#include <fstream>

#include "cli_common.hpp"

int main(int argc, char* argv[]) {
  if (argc < 9) {
    std::cout << "Usage: exe neuron1 batch1 layer1 neuron2 batch2 layer2 gpu_id dataset_dir" << std::endl;
    return 0;
  }
  const int n1 = std::atoi(argv[1]), b1 = std::atoi(argv[2]), l1 = std::atoi(argv[3]);
  const int n2 = std::atoi(argv[4]), b2 = std::atoi(argv[5]), l2 = std::atoi(argv[6]);
  const int dev = std::atoi(argv[7]);
  const std::string dir(argv[8]);
  cli::Flags f;
  if (!cli::parse_flags(argc, argv, 9, f)) return 2;
  std::cout << "-------------------------------------" << std::endl;
  std::cout << "[BEGIN]..." << std::endl;
  std::cout << "[Batch1] :" << b1 << std::endl << "[Neuron1] :" << n1 << std::endl << "[Layer1] :" << l1 << std::endl;
  std::cout << "-------------------------------------" << std::endl;
  std::cout << "[Batch2] :" << b2 << std::endl << "[Neuron2] :" << n2 << std::endl << "[Layer2] :" << l2 << std::endl;
  std::cout << "-------------------------------------" << std::endl;
  cli::Dataset d1, d2;
  cli::load_dataset(dir, n1, b1, l1, f.use_fuse, d1);
  cli::load_dataset(dir, n2, b2, l2, f.use_fuse, d2);
  if (l1 != l2 && !f.allow_uneven) {
    std::cerr << "Error: Two networks must have the same number of layers!" << std::endl;
    std::cout << "[END]..." << std::endl;
    return 0;
  }
  cospdnn_engine* eng = nullptr;
  int rc = cospdnn_engine_create(dev, &eng);
  if (rc != COSPDNN_OK) cli::die(rc);
  cospdnn_net* nets[2] = {nullptr, nullptr};
  rc = cli::make_net(eng, d1, 0, b1, f.bias_set ? f.bias : cli::bias_for(n1), f, &nets[0]);
  if (rc == COSPDNN_OK) rc = cli::make_net(eng, d2, 0, b2, f.bias_set ? f.bias : cli::bias_for(n2), f, &nets[1]);
  if (rc != COSPDNN_OK) cli::die(rc);
  std::vector<int64_t> c1(b1), c2(b2);
  cospdnn_result r[2];
  std::memset(r, 0, sizeof(r));
  r[0].categories = c1.data();
  r[1].categories = c2.data();
  const cospdnn_opts o = cli::run_opts(f);
  rc = cospdnn_run(nets, 2, &o, r);
  if (rc != COSPDNN_OK) cli::die(rc);
  c1.resize(r[0].n_categories);
  c2.resize(r[1].n_categories);
  std::cout << "categories = " << c1.size() << " and " << c2.size() << std::endl;
  // one concurrent run: both results report its time
  std::cout << "Total Kernel Exec Time = " << std::max(r[0].ms_total, r[1].ms_total) << "ms" << std::endl;
  if (!f.categories.empty()) {
    std::ofstream out(f.categories);
    for (int64_t id : c1) out << "1\t" << (id + 1) << "\n";
    for (int64_t id : c2) out << "2\t" << (id + 1) << "\n";
  }
  cospdnn_net_destroy(nets[0]);
  cospdnn_net_destroy(nets[1]);
  cospdnn_engine_destroy(eng);
  std::cout << "[END]..." << std::endl;
  return 0;
}
