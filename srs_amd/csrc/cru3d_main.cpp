// cru3d_main.cpp -- command-line driver with the reference's options
// (cpu/main.cpp:87-102): -param P.xml -mesh base [-out dir] [-start i] [-rank r],
// plus -layout gpu for states records in the GPU program's layout
// (gpu/OutputFuncs.cpp:277-284).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "srs_abi.h"

int main(int argc, char** argv) {
  const char* param = "TestSet.xml";  // DEFAULT_PARAMETER_FILE, cpu/parameters.h:105
  const char* mesh = "mesh";
  const char* out = nullptr;
  int rank = 0, start = 0, gpu_layout = 0;
  for (int i = 1; i < argc; i += 2) {
    if (i + 1 >= argc) {
      fprintf(stderr, "Error: option %s needs a value\n", argv[i]);
      return 1;
    }
    if (!strcmp(argv[i], "-rank")) rank = atoi(argv[i + 1]);
    else if (!strcmp(argv[i], "-start")) start = atoi(argv[i + 1]);
    else if (!strcmp(argv[i], "-param")) param = argv[i + 1];
    else if (!strcmp(argv[i], "-out")) out = argv[i + 1];
    else if (!strcmp(argv[i], "-mesh")) mesh = argv[i + 1];
    else if (!strcmp(argv[i], "-layout") && !strcmp(argv[i + 1], "gpu")) gpu_layout = 1;
    else if (!strcmp(argv[i], "-layout") && !strcmp(argv[i + 1], "cpu")) gpu_layout = 0;
    else {
      fprintf(stderr, "Error: unrecognized command line option: %s\n", argv[i]);
      return 1;
    }
  }
  double secs = 0;
  const int rc = srs_run_cru3d(param, mesh, out, rank, start, gpu_layout, &secs);
  if (rc != SRS_OK) {
    fprintf(stderr, "Error (%d): %s\n", rc, srs_last_error());
    return 1;
  }
  return 0;
}
