// The host stages of the reference launcher (src/launcher.cu:46-90,117-145)
// written against this library's drop-in C++ API: Hierarchical ->
// getRootIndex -> generateOps_GETRF -> h_ops_dag(tree, start, length) ->
// instructions_scheduler -> getLengths -> instructions_manager.  Compiled and
// run on the host by tests/test_dropin_cpp.py; prints one line of counts per
// window that the test compares with the C ABI's program of the same
// configuration.
#include <cstdio>
#include <cstdlib>

#include "h_ops/h_ops.h"
#include "instructions/instructions_manager.h"
#include "matrix/Hierarchical.h"
#include "rts/h_ops_dag.h"
#include "rts/scheduler.h"

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 2048, levels = argc > 2 ? std::atoi(argv[2]) : 3;
  const int rank = argc > 3 ? std::atoi(argv[3]) : 16, workers = argc > 4 ? std::atoi(argv[4]) : 8;
  const int kernel_size = argc > 5 ? std::atoi(argv[5]) : 0;
  Hierarchical* h = Hierarchical::hodlr(n, levels, rank);
  std::printf("matrix %d x %d\n", h->getNy_abs(), h->getNx_abs());
  dev_temp tmp_mngr;
  const h_index* root = h->getRootIndex();
  const h_ops_tree* tree = h->generateOps_GETRF(root, &tmp_mngr);
  delete root;
  const int iters = kernel_size <= 0 ? 1 : (tree->length() + kernel_size - 1) / kernel_size;
  for (int i = 0; i < iters; ++i) {
    const int start = i * (kernel_size <= 0 ? 0 : kernel_size);
    h_ops_dag dag(tree, start, kernel_size);
    instructions_scheduler* schedule = new instructions_scheduler(&dag, workers);
    int* lengths = schedule->getLengths();
    long long entries = 0;
    for (int w = 0; w < workers; ++w) entries += lengths[w];
    delete[] lengths;
    instructions_manager ins(workers, &dag, schedule, tree, start);
    std::printf("window %d insts %d entries %lld words %lld flops %lld density %.6f\n", i, dag.getLength(), entries,
                (long long)ins.totalWords(), (long long)dag.getFlops(), dag.dag_density());
    delete schedule;
  }
  std::printf("temps %lld\n", (long long)tmp_mngr.size());
  delete tree;
  delete h;
  return 0;
}
