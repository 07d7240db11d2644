// Determinant and wavefunction files through the drop-in C++ API (include/sbd_amd/sbd.h):
// sbd::LoadAlphaDets on a .txt or .bin file (words written flat, bit_length bits each) and,
// with a vector file, sbd::SaveMatrixFormWF as .bin and .txt.  Host code only: no GPU call.
// Test program: tests/test_shim_io.py compares its files with the reference's own.
//
//   shim_io <adets file> <bdets file> <norb> <bit_length> <out prefix> [vector file]
//   shim_io configs <dets .bin file> <norb> <bit_length> <out prefix>
//       sbd::LoadFromAlphaDets: <out>.config.u64 (the configurations' words, flat) and the
//       word count per configuration in the printed line
//   shim_io pack <words file> <norb> <bit_length> <out prefix>
//       the header's packing of bit_length-bit strings (the file: ceil(norb / bit_length)
//       uint64 words per string, flat) into the engine's 64-bit words: <out>.packed.u64
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "sbd_amd/sbd.h"

int main(int argc, char** argv) {
  int provided;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
  if (argc == 6 && std::string(argv[1]) == "configs") {
    std::vector<std::vector<size_t>> config;
    sbd::LoadFromAlphaDets(argv[2], config, std::stoul(argv[3]), std::stoul(argv[4]), MPI_COMM_WORLD);
    std::ofstream f(std::string(argv[5]) + ".config.u64", std::ios::binary);
    for (const auto& v : config) f.write(reinterpret_cast<const char*>(v.data()), sizeof(size_t) * v.size());
    std::printf("{\"op\":\"configs\",\"n\":%zu,\"words\":%zu}\n", config.size(),
                config.empty() ? size_t(0) : config[0].size());
    MPI_Finalize();
    return 0;
  }
  if (argc == 6 && std::string(argv[1]) == "pack") {
    const size_t norb = std::stoul(argv[3]), bl = std::stoul(argv[4]), nw = (norb + bl - 1) / bl;
    std::ifstream in(argv[2], std::ios::binary);
    std::vector<std::vector<size_t>> d;
    std::vector<size_t> w(nw);
    while (in.read(reinterpret_cast<char*>(w.data()), sizeof(size_t) * nw)) d.push_back(w);
    int rc = 0;
    try {
      const std::vector<uint64_t> p = sbd::amd_detail::pack(d, bl, norb);
      std::ofstream f(std::string(argv[5]) + ".packed.u64", std::ios::binary);
      f.write(reinterpret_cast<const char*>(p.data()), sizeof(uint64_t) * p.size());
      std::printf("{\"op\":\"pack\",\"n\":%zu,\"words\":%d}\n", d.size(), sbd::amd_detail::words_for(norb));
    } catch (const std::invalid_argument& e) {
      std::printf("{\"op\":\"pack\",\"error\":\"%s\"}\n", e.what());
      rc = 1;
    }
    MPI_Finalize();
    return rc;
  }
  if (argc < 6) {
    std::fprintf(stderr, "usage: shim_io <adets> <bdets> <norb> <bit_length> <out prefix> [vector] | configs <bin> <norb> <bit_length> <out>\n");
    return 2;
  }
  const size_t norb = std::stoul(argv[3]), bit_length = std::stoul(argv[4]);
  const std::string out = argv[5];
  std::vector<std::vector<size_t>> adet, bdet;
  sbd::LoadAlphaDets(argv[1], adet, bit_length, norb);
  sbd::LoadAlphaDets(argv[2], bdet, bit_length, norb);
  auto dump = [](const std::string& path, const std::vector<std::vector<size_t>>& d) {
    std::ofstream f(path, std::ios::binary);
    for (const auto& v : d) f.write(reinterpret_cast<const char*>(v.data()), sizeof(size_t) * v.size());
  };
  dump(out + ".adet.u64", adet);
  dump(out + ".bdet.u64", bdet);
  if (argc > 6) {
    std::vector<double> W(adet.size() * bdet.size());
    std::ifstream f(argv[6], std::ios::binary);
    f.read(reinterpret_cast<char*>(W.data()), static_cast<std::streamsize>(8 * W.size()));
    if (!f) return 3;
    MPI_Comm h_comm, b_comm, t_comm;
    sbd::TaskCommunicator(MPI_COMM_WORLD, 1, 1, 1, 1, h_comm, b_comm, t_comm);
    sbd::SaveMatrixFormWF(out + ".wf.bin", adet, bdet, 1, 1, h_comm, b_comm, t_comm, W);
    sbd::SaveMatrixFormWF(out + ".wf.txt", adet, bdet, 1, 1, h_comm, b_comm, t_comm, W);
  }
  std::printf("{\"op\":\"files\",\"na\":%zu,\"nb\":%zu}\n", adet.size(), bdet.size());
  MPI_Finalize();
  return 0;
}
