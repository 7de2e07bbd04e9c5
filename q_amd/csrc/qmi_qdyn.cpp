// qmi-qdyn: command-line form of the engine, a drop-in for `qdyn <input.inp>`.
//
// The reference reads the input file named by its last argument
// (src/q6/md.f90:2568-2577), writes its log to stdout (write_out,
// md.f90:15292-15438) and ends with "terminated normally" (qdyn.f90:224-233)
// or, on a fatal error, "qdyn terminated abnormally" (die, md.f90:400-455).
// This command keeps that contract for one input, so `qmi-qdyn x.inp > x.log`
// replaces `qdyn x.inp > x.log`; given several inputs it advances them as one
// batch on the GPU and each window logs to <input>.log.
//
//   qmi-qdyn [--gpu N] [--seed S] input.inp [more.inp ...]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qmi.h"

static int die(const char* what) {
  std::fprintf(stderr, ">>> ERROR: %s: %s\n", what, qmi_last_error());
  std::fprintf(stdout, "qdyn terminated abnormally\n");
  return 255;
}

int main(int argc, char** argv) {
  int gpu = 0;
  long long seed = -1;
  std::vector<std::string> inputs;
  for (int a = 1; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--gpu") && a + 1 < argc) {
      gpu = std::atoi(argv[++a]);
    } else if (!std::strcmp(argv[a], "--seed") && a + 1 < argc) {
      seed = std::atoll(argv[++a]);
    } else if (!std::strcmp(argv[a], "-h") || !std::strcmp(argv[a], "--help")) {
      std::printf("usage: qmi-qdyn [--gpu N] [--seed S] input.inp [more.inp ...]\n");
      return 0;
    } else if (argv[a][0] == '-' && argv[a][1] == '-') {
      std::fprintf(stderr, "unknown option %s\n", argv[a]);
      return 2;
    } else {
      inputs.push_back(argv[a]);
    }
  }
  if (inputs.empty()) {
    std::fprintf(stderr, "usage: qmi-qdyn [--gpu N] [--seed S] input.inp [more.inp ...]\n");
    return 2;
  }
  qmi_engine* e = nullptr;
  if (qmi_engine_create(gpu, &e)) return die("engine");
  const long long longest = 1LL << 40;
  for (const std::string& inp : inputs) {
    int w = -1;
    if (qmi_window_add(e, inp.c_str(), seed, &w)) {
      std::fprintf(stderr, "in %s\n", inp.c_str());
      qmi_engine_destroy(e);
      return die("reading input");
    }
    if (inputs.size() == 1 && qmi_window_set_log(e, w, "-")) return die("log");
  }
  if (qmi_prepare(e)) {
    qmi_engine_destroy(e);
    return die("preparing the GPU run");
  }
  // every window stops at its own [MD] steps; the run ends when all have
  if (qmi_run(e, longest, 1)) {
    qmi_engine_destroy(e);
    return die("running");
  }
  qmi_engine_destroy(e);
  std::fflush(stdout);
  std::printf(" qmi-qdyn (qdyn 6.0.1 files) terminated normally.\n");
  for (int i = 0; i < 80; ++i) std::putchar('=');
  std::putchar('\n');
  return 0;
}
