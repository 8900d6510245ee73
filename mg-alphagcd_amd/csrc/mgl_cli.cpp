// mgl_cli: command-line front end of libmgl.so with the MG_GCD flags
// (src/main.cu:20-26) and its stdout layout (graph banner
// src/graph/host_graph.cpp:9-13, per-phase loop table src/louvain/louvain.cu:
// 1536-1538 and :1654, phase and total times :1670-1703).
//
//   mgl_cli -path_of_graph G.mtx [-threshold 1e-6] [-max_phases 1000]
//           [-max_iter 500] [-version 2] [-precision auto|exact|fp64|fp32]
//           [-honour_max_iter 0|1] [-out labels.txt] [-device 0]
//           [-dump PREFIX [-dump_binary 0|1]]   levels, dendrogram, coarse graphs
//   mgl_cli -rmat SCALE,EF,SEED [...]      synthetic R-MAT (0.57,0.19,0.19,0.05)
//   mgl_cli -sbm N,COMMUNITIES,SEED [...]  synthetic planted partition
//
// As in the reference, -max_iter is accepted but not applied unless
// -honour_max_iter 1 is given.
#include "../../include/mgl.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

static const char* arg(int argc, char** argv, const char* key, const char* def) {
  for (int i = 1; i + 1 < argc; ++i)
    if (strcmp(argv[i], key) == 0) return argv[i + 1];
  return def;
}

int main(int argc, char** argv) {
  const std::string path = arg(argc, argv, "-path_of_graph", "");
  const double thr = atof(arg(argc, argv, "-threshold", "1e-6"));
  const int max_iter = atoi(arg(argc, argv, "-max_iter", "500"));
  const int max_phases = atoi(arg(argc, argv, "-max_phases", "1000"));
  const int version = atoi(arg(argc, argv, "-version", "2"));
  const int honour = atoi(arg(argc, argv, "-honour_max_iter", "0"));
  const std::string prec = arg(argc, argv, "-precision", "auto");
  const char* out = arg(argc, argv, "-out", nullptr);
  const char* dump = arg(argc, argv, "-dump", nullptr);
  const int dump_binary = atoi(arg(argc, argv, "-dump_binary", "0"));
  const int device = atoi(arg(argc, argv, "-device", "0"));
  const char* rmat = arg(argc, argv, "-rmat", nullptr);
  const char* sbm = arg(argc, argv, "-sbm", nullptr);

  mgl_host_csr h;
  memset(&h, 0, sizeof(h));
  mgl_status st;
  std::string name = path;
  if (rmat) {
    int scale = 16, seed = 1;
    double ef = 16;
    sscanf(rmat, "%d,%lf,%d", &scale, &ef, &seed);
    st = mgl_gen_rmat(scale, ef, 0.57, 0.19, 0.19, (uint64_t)seed, 0, 0, &h);
    name = std::string("rmat:") + rmat;
  } else if (sbm) {
    long long n = 1 << 20, k = 1024;
    int seed = 2;
    sscanf(sbm, "%lld,%lld,%d", &n, &k, &seed);
    st = mgl_gen_sbm(n, k, 2.1, 64, 24, 4, (uint64_t)seed, 1, &h);
    name = std::string("sbm:") + sbm;
  } else if (!path.empty()) {
    st = mgl_read_mtx(path.c_str(), &h);
  } else {
    fprintf(stderr, "usage: %s -path_of_graph G.mtx | -rmat S,EF,SEED | -sbm N,K,SEED [flags]\n",
            argv[0]);
    return 2;
  }
  if (st != MGL_OK) {
    fprintf(stderr, "cannot load graph %s (status %d)\n", name.c_str(), (int)st);
    return 1;
  }
  printf("%s\n", std::string(75, '-').c_str());
  printf("%25s%25s%25s\n", "Input graph", "Num. vertices (n)", "Num. edges (M)");
  printf("%s\n", std::string(75, '-').c_str());
  printf("%25s%25lld%25lld\n", name.c_str(), (long long)h.n, (long long)h.nnz);
  printf("%s\n", std::string(75, '-').c_str());

  mgl_ctx* ctx = nullptr;
  if (mgl_create(&ctx, device) != MGL_OK) {
    fprintf(stderr, "cannot open device %d\n", device);
    mgl_csr_free(&h);
    return 1;
  }
  mgl_opts o;
  mgl_opts_default(&o);
  o.version = version;
  o.threshold = thr;
  o.max_phases = max_phases;
  o.max_iter = honour ? max_iter : 0;
  o.precision = prec == "exact" ? MGL_PREC_EXACT
              : prec == "fp64"  ? MGL_PREC_FP64
              : prec == "fp32"  ? MGL_PREC_FP32_HASH
                                : MGL_PREC_AUTO;
  if (dump) o.keep_levels = o.keep_graphs = 1;
  mgl_csr g = {h.n, h.nnz, h.offsets, h.cols, h.w, h.w64};
  mgl_result* r = nullptr;
  st = mgl_louvain(ctx, &g, &o, &r);
  if (st != MGL_OK) {
    fprintf(stderr, "louvain failed: %s\n", mgl_last_error(ctx));
    mgl_destroy(ctx);
    mgl_csr_free(&h);
    return 1;
  }
  double loops = 0, coarse = 0;
  const int L = mgl_result_levels(r);
  for (int l = 0; l < L; ++l) {
    printf("----------------------------------------\n");
    printf("Phase %d\n", l);
    printf("| %-10s | %-10s | %-10s | %-10s | %-10s |\n", "Loop", "Q", "dQ", "time(s)", "time(ms)");
    printf("|------------|------------|------------|------------|------------|\n");
    const int it = mgl_result_level_iters(r, l);
    const double* qt = mgl_result_level_qtraj(r, l);
    const double* ts = mgl_result_level_iter_seconds(r, l);
    double cs = 0;
    const double ls = mgl_result_level_seconds(r, l, &cs);
    // loop 0: the singleton start (louvain.cu:1536-1538)
    double cur = mgl_result_level_q0(r, l);
    printf("| %-10d | %-10f | %-10f | %-10f | %-10f |\n", 0, cur, 0., 0., 0.);
    double loop_s = 0;
    for (int i = 0; i < it; ++i) {
      // dQ = new_Q - cur_Q, zero after a rejected proposal (louvain.cu:1642-1654)
      const double t = ts ? ts[i] : (it > 0 ? ls / it : 0);
      printf("| %-10d | %-10f | %-10f | %-10f | %-10f |\n", i + 1, qt[i], qt[i] - cur, t, t * 1000);
      cur = qt[i];
      loop_s += t;
    }
    const double per = it > 0 ? loop_s / it : 0;
    printf("Main loop total execution time: %f s, %f ms\n", ls, ls * 1000);
    printf("The average execution time per loop: %f s, %f ms\n", per, per * 1000);
    if (l + 1 < L) printf("Coarsen graph execution time: %f s, %f ms\n", cs, cs * 1000);
    loops += ls;
    coarse += cs;
  }
  printf("----------------------------------------\n");
  printf("Total time for clustering   : %f s\n", loops);
  printf("Total time for coarsen graph: %f s\n", coarse);
  printf("TOTAL TIME                  : %f s\n", mgl_result_seconds(r));
  printf("Final modularity            : %.12f (levels %d)\n", mgl_result_final_q(r), L);
  if (out) {
    if (mgl_write_assign(out, mgl_result_final_assign(r), h.n) != MGL_OK)
      fprintf(stderr, "cannot write %s\n", out);
  }
  if (dump && mgl_write_result(r, dump, dump_binary ? 1 : 0) != MGL_OK)
    fprintf(stderr, "cannot write the result files %s.*\n", dump);
  mgl_result_free(r);
  mgl_destroy(ctx);
  mgl_csr_free(&h);
  return 0;
}
