// driver_main.cpp -- command-line entry of the RTM driver.
//
//   rtm_driver [2D_Real_RVSP_RTM.txt] [--workers N] [--batch B] [--retries R]
//              [--strips f32|f16|bf16] [--no-resume] [--template SGY_Model.sgy] [-v]
//
// Like the reference's ./RTM, the parameter file defaults to
// 2D_Real_RVSP_RTM.txt in the working directory.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../../include/rtm_hip.h"
#include "../../../include/rtm_host.h"

int main(int argc, char** argv) {
  const char* cfg = "2D_Real_RVSP_RTM.txt";
  rtmh_run_opts o{1, 8, 1, RTM_STRIP_F32, 1, 0, nullptr};
  for (int i = 1; i < argc; ++i) {
    auto next = [&](const char* what) -> const char* {
      if (i + 1 >= argc) {
        fprintf(stderr, "missing value for %s\n", what);
        exit(2);
      }
      return argv[++i];
    };
    if (!strcmp(argv[i], "--workers")) o.workers = atoi(next("--workers"));
    else if (!strcmp(argv[i], "--batch")) o.batch = atoi(next("--batch"));
    else if (!strcmp(argv[i], "--retries")) o.retries = atoi(next("--retries"));
    else if (!strcmp(argv[i], "--no-resume")) o.resume = 0;
    else if (!strcmp(argv[i], "--template")) o.sgy_template = next("--template");
    else if (!strcmp(argv[i], "-v")) o.verbose = 1;
    else if (!strcmp(argv[i], "--strips")) {
      const char* s = next("--strips");
      o.strip_format = !strcmp(s, "f16") ? RTM_STRIP_F16 : !strcmp(s, "bf16") ? RTM_STRIP_BF16 : RTM_STRIP_F32;
    } else if (argv[i][0] == '-') {
      fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    } else {
      cfg = argv[i];
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  int rc = rtmh_run(cfg, &o);
  double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc) {
    fprintf(stderr, "rtm_driver: failed (%d): %s\n", rc, rtmh_last_error());
    return 1;
  }
  printf("%f seconds\n", s);
  return 0;
}
