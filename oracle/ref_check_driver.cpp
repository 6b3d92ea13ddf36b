// Reference checker build driver (test infrastructure only).
//
// Compiles test-results/cluster-check.cpp in place (included by path, never
// copied), renames its main() (which reads two hard-coded files of exactly
// DATASET_SIZE labels) and instead feeds DetermineErrorTwoClusterResults
// (cluster-check.cpp:41-199) the two label files named on the command line.
#define main cluster_check_unused_main
#include "cluster-check.cpp"
#undef main

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  std::vector<int> a, b;
  for (int f = 0; f < 2; ++f) {
    FILE* fp = fopen(argv[1 + f], "r");
    if (!fp) return 1;
    int v;
    while (fscanf(fp, "%d", &v) == 1) (f ? b : a).push_back(v);
    fclose(fp);
  }
  DetermineErrorTwoClusterResults(&a, &b);
  printf("\n");
  return 0;
}
