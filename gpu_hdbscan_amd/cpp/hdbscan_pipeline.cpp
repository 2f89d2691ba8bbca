// hdbscan_pipeline -- C++ driver of the hot path through include/gpu_hdbscan.hpp,
// in the call order of the reference driver HDBSCAN/_new.cu:149-357
// (kNNG -> enhanced kNNG -> MST -> edge sort -> labels), with wall-clock
// stage timers (the reference prints clock() CPU seconds).
//
//   hdbscan_pipeline <input.fvecs> <mpts> <shards> <labels_out.ivecs> [iterations]
//
// shards > 1 builds the kNN graph with the blocked all-pairs schedule of
// GenLargeKNNGraph; shards == 1 runs one NN-Descent over all points.
// .fvecs / .ivecs rows are (int32 dim, dim values), as in tools/filetool.hpp:95-242.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "../../include/gpu_hdbscan.hpp"

static std::vector<float> read_fvecs(const char* path, int* n, int* dim) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  std::vector<float> data;
  int d = 0, rows = 0;
  while (true) {
    int dd;
    if (!f.read(reinterpret_cast<char*>(&dd), 4)) break;
    if (rows == 0) d = dd;
    if (dd != d || d <= 0) throw std::runtime_error("ragged or empty .fvecs row");
    size_t off = data.size();
    data.resize(off + d);
    if (!f.read(reinterpret_cast<char*>(data.data() + off), d * sizeof(float))) throw std::runtime_error("short row");
    ++rows;
  }
  *n = rows;
  *dim = d;
  return data;
}

static void write_ivecs(const char* path, const std::vector<int>& v) {
  std::ofstream f(path, std::ios::binary);
  for (int x : v) {
    const int one = 1;
    f.write(reinterpret_cast<const char*>(&one), 4);
    f.write(reinterpret_cast<const char*>(&x), 4);
  }
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s <input.fvecs> <mpts> <shards> <labels_out.ivecs> [iterations]\n", argv[0]);
    return 1;
  }
  const int mpts = std::atoi(argv[2]);
  const int shards = std::atoi(argv[3]);
  const int iterations = argc > 5 ? std::atoi(argv[5]) : 6;
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
  try {
    int n = 0, dim = 0;
    std::vector<float> X = read_fvecs(argv[1], &n, &dim);
    std::printf("NUM VALUES SET TO %d, dim %d, mpts %d, shards %d\n", n, dim, mpts, shards);
    hdb::Context ctx(0);
    hdb::DevBuf<float> Xd(ctx, X.size());
    ctx.h2d(Xd.p, X.data(), X.size() * sizeof(float));
    auto t = clk::now();
    NNDElement* knn = nullptr;
    if (shards > 1) {
      knn = static_cast<NNDElement*>(ctx.alloc(static_cast<size_t>(n) * 32 * sizeof(NNDElement)));
      GenLargeKNNGraph(ctx, Xd.p, n, dim, shards, knn);
    } else {
      gpuknn::NNDescent(ctx, &knn, Xd.p, n, dim, iterations, true);
    }
    hdb_synchronize(ctx.get());
    std::printf("TEMPO KNNG: %lf\n", secs(t));
    t = clk::now();
    ECLgraph g = buildEnhancedKNNG(ctx, Xd.p, knn, n, dim, mpts);
    std::printf("TEMPO ENHANCED KNNG: %lf (edges %ld)\n", secs(t), g.edges);
    t = clk::now();
    long comps = 0;
    std::vector<MSTedge> mst = buildMST(ctx, g, &comps);
    sort_edges(ctx, mst);
    std::printf("TEMPO MST: %lf (%zu edges, %ld components)\n", secs(t), mst.size(), comps);
    freeECLgraph(g);
    ctx.free(knn);
    if (comps != 1) {
      std::printf("MST montada incorretamente: %ld components\n", comps);
      return 3;
    }
    t = clk::now();
    int nc = 0;
    std::vector<int> labels = get_clusters(mst, n, mpts, &nc);
    std::printf("TEMPO LABELS: %lf (clusters %d)\n", secs(t), nc);
    write_ivecs(argv[4], labels);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  return 0;
}
