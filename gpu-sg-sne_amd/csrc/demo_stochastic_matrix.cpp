// demo_stochastic_matrix -- the SG-t-SNE-Pi / HYB command-line program
// (HY/demo_stochastic_matrix.cpp:19-99) over this engine:
//
//   sgtsne_demo_stochastic_matrix [-l lambda] [-d dims] [-a alpha] [-m maxIter]
//                                 [-e earlyIter] [-h h] [-p np] file.mtx
//
// Loads a sparse stochastic graph in Matrix Market format (1-based "row col
// value" triplets), runs SG-t-SNE through the HYB interface (FP64, AoS,
// column-stochastic CSC, made stochastic first) and writes embedding.txt, one
// line per point.  As in the reference's sgtsne (HY/sgtsne.cpp:153-168), the
// start embedding is read from yin2m.txt (n lines of d values) when that file
// exists; otherwise the reference CUDA path's random start is used.
//
// SGTSNE_DEMO_DRY_RUN=1 in the environment stops after the options and the
// graph are read and prints them, without touching the device.
#include "sgtsne.hpp"

#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
  int opt;
  sgtsne_hyb::tsneparams params;
  std::string filename;

  if ((argc <= 1) || (argv[argc - 1] == nullptr) || (argv[argc - 1][0] == '-')) {
    std::cerr << "No filename provided!" << std::endl;
    return 1;
  }
  filename = argv[argc - 1];

  opterr = 0;
  while ((opt = getopt(argc, argv, "l:d:a:m:e:h:p:")) != -1) {
    switch (opt) {
      case 'l': std::sscanf(optarg, "%lf", &params.lambda); break;
      case 'd': params.d = std::atoi(optarg); break;
      case 'm': params.maxIter = std::atoi(optarg); break;
      case 'e': params.earlyIter = std::atoi(optarg); break;
      case 'p': params.np = std::atoi(optarg); break;
      case 'a': std::sscanf(optarg, "%lf", &params.alpha); break;
      case 'h': std::sscanf(optarg, "%lf", &params.h); break;
      case '?': std::cerr << "Unknown option: '" << char(optopt) << "'!" << std::endl; break;
    }
  }

  // ----- load the stochastic graph (compressed columns)
  int m = 0, ncol = 0;
  int64_t nnz = 0;
  int32_t *colptr = nullptr, *row = nullptr;
  double *val = nullptr;
  sgtsne_mi355x::check(sgtsne_read_P_mtx(filename.c_str(), &m, &ncol, &nnz, &colptr, &row, &val), "buildPFromMTX");
  params.n = m;
  if (const char *dry = std::getenv("SGTSNE_DEMO_DRY_RUN"); dry && dry[0] == '1') {
    std::printf("lambda=%g d=%d alpha=%g maxIter=%d earlyIter=%d h=%g np=%d n=%d\n", params.lambda, params.d,
                params.alpha, params.maxIter, params.earlyIter, params.h, params.np, params.n);
    double sum = 0.0;
    for (int64_t t = 0; t < nnz; ++t) sum += val[t];
    std::printf("m=%d n=%d nnz=%lld sum=%.17g\n", m, ncol, (long long)nnz, sum);
    sgtsne_free(colptr);
    sgtsne_free(row);
    sgtsne_free(val);
    return 0;
  }
  if (m != ncol) {
    std::cerr << "the graph must be square (" << m << " x " << ncol << ")" << std::endl;
    return 1;
  }
  sgtsne_hyb::sparse_matrix P;
  P.m = m;
  P.n = ncol;
  P.nnz = (int)nnz;
  P.col = reinterpret_cast<uint32_t *>(colptr);  // int32 offsets and ids are < 2^31: same bits as uint32
  P.row = reinterpret_cast<uint32_t *>(row);
  P.val = val;

  // ----- start embedding
  std::vector<double> yin;
  {
    std::ifstream yinf("yin2m.txt");
    if (yinf.is_open()) {
      yin.resize((size_t)params.n * params.d);
      for (auto &v : yin) yinf >> v;
      if (!yinf) {
        std::cerr << "yin2m.txt holds fewer than " << params.n << " x " << params.d << " values" << std::endl;
        return 1;
      }
    }
  }
  double *y = sgtsne_hyb::sgtsne(P, params, yin.empty() ? nullptr : yin.data());
  sgtsne_mi355x::check(sgtsne_write_embedding_text_f64("embedding.txt", y, params.n, params.d, SGTSNE_LAYOUT_AOS),
                       "extractEmbeddingText");
  std::cout << std::endl
            << " If gnuplot is installed, issue" << std::endl
            << std::endl
            << "   gnuplot -e 'set size square; plot \"embedding.txt\" with dots'" << std::endl
            << std::endl
            << " to visualize the embedding" << std::endl
            << std::endl;
  std::free(y);
  sgtsne_free(colptr);
  sgtsne_free(row);
  sgtsne_free(val);
  return 0;
}
