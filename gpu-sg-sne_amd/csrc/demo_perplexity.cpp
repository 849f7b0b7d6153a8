// demo_perplexity -- the CUDA path's command-line program
// (CU/testing_modules/demo_perplexity.cu:33-226) over this engine:
//
//   sgtsne_demo_perplexity [-d dims] [-a alpha] [-m maxIter] [-e earlyIter]
//                          [-h h] [-p np] [-u perplexity] [-f computeError]
//                          [-g ng] [-s sim] [-n points] [-t format] file
//
// Reads the points (text "n dim" header, then the values) or, with -s C > 0,
// draws the reference's synthetic data (C clusters, 50-D); finds the
// ceil(3u) nearest neighbours (exact here; FLANN in the reference), divides
// the squared distances by their maximum, equalises perplexity, runs
// SG-t-SNE and writes sgtsneEmbedding.txt (one line per point) and, when it
// exists already, appends the timing report to timeInfo.txt.  As in the
// reference, with -s 0 the start embedding is read from yin2m.txt ([d][N]
// order, n lines of d values) when that file exists; otherwise the reference's
// own random start is used.
//
// SGTSNE_DEMO_DRY_RUN=1 in the environment stops after the options and the
// input are read and prints them, without touching the device.
#include "sgtsne.hpp"

#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
  int opt;
  coord u = 30;
  tsneparams params;
  std::string filename;

  if ((argc <= 1) || (argv[argc - 1] == nullptr) || (argv[argc - 1][0] == '-')) {
    std::cerr << "No filename provided!" << std::endl;
    return 1;
  }
  filename = argv[argc - 1];

  opterr = 0;
  while ((opt = getopt(argc, argv, "d:a:m:e:h:p:u:f:g:s:n:t:")) != -1) {
    switch (opt) {
      case 'd': params.d = std::atoi(optarg); break;
      case 'm': params.maxIter = std::atoi(optarg); break;
      case 'e': params.earlyIter = std::atoi(optarg); break;
      case 'p': params.np = std::atoi(optarg); break;
      case 'u': std::sscanf(optarg, "%f", &u); break;
      case 'a': std::sscanf(optarg, "%f", &params.alpha); break;
      case 'h': std::sscanf(optarg, "%f", &params.h); break;
      case 'f': std::sscanf(optarg, "%d", &params.ComputeError); break;
      case 'g': std::sscanf(optarg, "%d", &params.ng); break;
      case 's': std::sscanf(optarg, "%d", &params.sim); break;
      case 'n': std::sscanf(optarg, "%d", &params.n); break;
      case 't': std::sscanf(optarg, "%d", &params.format); break;
      case '?': std::cerr << "Unknown option: '" << char(optopt) << "'!" << std::endl; break;
    }
  }
  const int nn = (int)std::ceil(u * 3);
  params.lambda = 1;

  // ----- points
  float *X = nullptr;
  int n = 0, dim = 0;
  if (params.sim == 0) {
    n = params.n;
    sgtsne_mi355x::check(sgtsne_read_X_text(filename.c_str(), &n, &dim, &X), "readXfromMTX");
  } else {
    n = params.n;
    dim = 50;
    X = static_cast<float *>(std::malloc(sizeof(float) * (size_t)n * dim));
    if (!X) sgtsne_mi355x::fail(SGTSNE_ERR_NOMEM, "loadSimulationData");
    sgtsne_mi355x::check(sgtsne_simulation_data(n, dim, params.sim, X), "loadSimulationData");
  }
  params.n = n;
  if (const char *dry = std::getenv("SGTSNE_DEMO_DRY_RUN"); dry && dry[0] == '1') {
    std::printf("d=%d alpha=%g maxIter=%d earlyIter=%d h=%g np=%d u=%g ComputeError=%d ng=%d sim=%d n=%d format=%d\n",
                params.d, (double)params.alpha, params.maxIter, params.earlyIter, (double)params.h, params.np,
                (double)u, params.ComputeError, params.ng, params.sim, params.n, params.format);
    std::printf("points=%d dim=%d neighbours=%d file=%s\n", n, dim, nn, filename.c_str());
    double sum = 0.0;
    for (size_t t = 0; t < (size_t)n * dim; ++t) sum += X[t];
    std::printf("sum=%.17g\n", sum);
    sgtsne_free(X);
    return 0;
  }
  std::cout << "Number of points " << n << " dimensions " << dim << "\n";
  std::cout << "Running k-nearest neighbor search for " << nn << " neighbors..." << std::flush;

  // ----- kNN (nn + 1 with the point itself first), squared distances
  std::vector<int> I((size_t)n * (nn + 1));
  std::vector<double> Dd((size_t)n * (nn + 1));
  sgtsne_mi355x::check(sgtsne_knn(X, n, dim, nn, SGTSNE_KNN_AUTO, I.data(), Dd.data(), nullptr), "allKNNsearch");
  sgtsne_free(X);
  // the reference's float distances, max-normalised (demo_perplexity.cu:155-157)
  std::vector<coord> D(Dd.size());
  coord mx = 0;
  for (size_t t = 0; t < D.size(); ++t) {
    D[t] = (coord)Dd[t];
    if (D[t] > mx) mx = D[t];
  }
  if (mx > 0)
    for (auto &v : D) v /= mx;
  std::cout << "DONE" << std::endl;
  sparse_matrix<coord> P = perplexityEqualization<coord>(I.data(), D.data(), n, nn, u);
  const int d = params.d;

  // ----- start embedding
  std::vector<coord> yin;
  if (params.sim == 0) {
    std::ifstream yinf("yin2m.txt");
    if (yinf.is_open()) {
      yin.resize((size_t)n * d);
      for (int i = 0; i < n; i++)
        for (int j = 0; j < d; j++) yinf >> yin[i + (size_t)j * n];
      if (!yinf) {
        std::cerr << "yin2m.txt holds fewer than " << n << " x " << d << " values" << std::endl;
        return 1;
      }
    }
  }
  std::vector<double> timeInfo((size_t)7 * std::max(params.maxIter, 0) + 3, 0.0);
  const auto t1 = std::chrono::steady_clock::now();
  coord *y = sgtsneCUDA<coord>(&P, params, yin.empty() ? nullptr : yin.data(), timeInfo.data());
  const auto t2 = std::chrono::steady_clock::now();
  const double elapsed = std::chrono::duration<double, std::milli>(t2 - t1).count();
  std::printf("Total time=%lf\n", elapsed);
  timeInfo[(size_t)7 * params.maxIter + 2] = elapsed;
  sgtsne_mi355x::check(sgtsne_write_time_info("timeInfo.txt", timeInfo.data(), n, d, params.maxIter),
                       "ExtractTimeInfo");
  sgtsne_mi355x::check(sgtsne_write_embedding_text("sgtsneEmbedding.txt", y, params.n, d, SGTSNE_LAYOUT_SOA),
                       "extractEmbeddingTextT");
  std::free(y);
  std::free(P.row);
  std::free(P.col);
  std::free(P.val);
  return 0;
}
